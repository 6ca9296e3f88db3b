// Shader-independent kernels: triangle setup, prefix scans and the
// order-preserving tile binning (emit + stable LSD radix sort by tile).
//
// Binning must keep submission order inside every tile: blending and the
// LESS depth test are order dependent, and the reference draws primitives
// strictly in order (CPVulkan/CommandBuffer.Draw.cpp:1526-1593).  Pairs
// (tile, primitive) are emitted in primitive order and then sorted by tile
// with a stable radix sort whose per-wave ranks come from wavefront ballots.
#pragma once
#pragma clang fp contract(off)

#include "raster_device.h"

namespace cpv {

// --------------------------------------------------------------- setup

struct SetupArgs {
    const DrawDesc* draws;
    const uint32_t* block_draw;   // setup-grid block -> draw
    const TargetParams* targets;
    const float* vout;
    TriSetup* recs;
    LineExt* ext;                 // per-primitive line extension (launches with lines)
    uint32_t* counts;             // tiles overlapped per primitive (0 = dropped)
    uint32_t stride;
    int32_t tile_shift;
};

// EdgeFunction(a, b, c) with c the pixel centre, stored as a, (b.y - a.y), (b.x - a.x).
__device__ __forceinline__ void put_edge(TriSetup& r, int k, const float* a, const float* b)
{
    r.ax[k] = a[0];
    r.ay[k] = a[1];
    r.dy[k] = b[1] - a[1];
    r.dx[k] = b[0] - a[0];
}

// Interval holding the viewport-transformed depth dt of every pixel the
// triangle covers, for exact hierarchical depth culling.
//
// A covered pixel has fp32 edge values e_i >= 0 and weights w_i = e_i / A
// (A the fp32 area).  Each e_i differs from the exact edge function of the
// same fp32 inputs by at most 5u * BW * BH * 2 (u = 2^-24, BW x BH the
// bounding box in NDC grown by the pixels the truncated box can add), the
// exact edge functions sum to the exact area, and A differs from that by
// at most the same bound.  So S = sum w_i lies in [1 - rho, 1 + rho] with
// rho = 41u * kappa (+u), kappa = BW * BH / A; rho = 64u * kappa + 8u is used.
// With w_i >= 0 the real sum of z_i * w_i lies between zmin * S and zmax * S,
// and the fp32 evaluation adds at most 4u * max|z| * S.  The bounds are
// widened by 16u * max|z| to absorb their own rounding, then mapped through
// the monotone viewport transform.  Degenerate, huge or non-finite cases
// get the whole real line.
__device__ __forceinline__ void depth_bounds(TriSetup& r, const float P[3][4], const DrawDesc& d)
{
    const float inf = __builtin_inff();
    r.dt_lo = -inf;
    r.dt_hi = inf;
    const float z0 = P[0][2], z1 = P[1][2], z2 = P[2][2];
    const float minx = fminf(P[0][0], fminf(P[1][0], P[2][0])), maxx = fmaxf(P[0][0], fmaxf(P[1][0], P[2][0]));
    const float miny = fminf(P[0][1], fminf(P[1][1], P[2][1])), maxy = fmaxf(P[0][1], fmaxf(P[1][1], P[2][1]));
    const float bw = (maxx - minx) + 8.0f / d.vp_w, bh = (maxy - miny) + 8.0f / d.vp_h;
    const float area = r.area;
    if (!(area > 0.0f) || !__builtin_isfinite(area) || !__builtin_isfinite(bw * bh) || !__builtin_isfinite(z0) ||
        !__builtin_isfinite(z1) || !__builtin_isfinite(z2) || fmaxf(fabsf(minx), fmaxf(fabsf(maxx), fmaxf(fabsf(miny), fabsf(maxy)))) > 1e15f)
        return;
    const float u = 0x1p-24f;
    const float rho = 64.0f * u * (bw * bh / area) + 8.0f * u;
    if (!(rho < 0.25f)) return;
    const float zmin = fminf(z0, fminf(z1, z2)), zmax = fmaxf(z0, fmaxf(z1, z2));
    const float zabs = fmaxf(fabsf(zmin), fabsf(zmax));
    const float lo = (zmin >= 0.0f ? zmin * (1.0f - rho) : zmin * (1.0f + rho)) - 16.0f * u * zabs * (1.0f + rho) - 0x1p-120f;
    const float hi = (zmax >= 0.0f ? zmax * (1.0f + rho) : zmax * (1.0f - rho)) + 16.0f * u * zabs * (1.0f + rho) + 0x1p-120f;
    const float sc = d.vp_max_depth - d.vp_min_depth;
    const float tl = sc * lo + d.vp_min_depth, th = sc * hi + d.vp_min_depth;
    if (!__builtin_isfinite(tl) || !__builtin_isfinite(th)) return;
    r.dt_lo = fminf(tl, th);
    r.dt_hi = fmaxf(tl, th);
}

// Vertex-output slot of assembled vertex `raw` of a draw (false: outside the
// bound vertex buffer).
__device__ __forceinline__ bool vertex_slot(const DrawDesc& d, uint32_t raw, uint32_t* slot)
{
    int64_t s = raw;
    if (d.vmode == kSlotVertexId) {
        uint32_t idx;
        if (d.index_size == 1) idx = ((const uint8_t*)d.index_addr)[d.first + raw];
        else if (d.index_size == 2) idx = ((const uint16_t*)d.index_addr)[d.first + raw];
        else idx = ((const uint32_t*)d.index_addr)[d.first + raw];
        s = (int64_t)d.vertex_offset + (int64_t)idx;
    }
    *slot = d.vout_base + (uint32_t)s;
    return s >= 0 && s < (int64_t)d.vout_count;
}

__device__ __forceinline__ void clip_to_ndc(const float* vout, uint32_t stride, uint32_t slot, float P[4])
{
    const float4 q = *(const float4*)(vout + (uint64_t)slot * stride);
    P[0] = q.x / q.w;
    P[1] = q.y / q.w;
    P[2] = q.z / q.w;
    P[3] = q.w;
}

__device__ __forceinline__ void store_record(TriSetup* dst_rec, const TriSetup& r)
{
    uint4* dst = (uint4*)dst_rec;
    const uint4* src = (const uint4*)&r;
#pragma unroll
    for (int q = 0; q < 8; q++) dst[q] = src[q];
}

__device__ __forceinline__ uint32_t tiles_of(int32_t x0, int32_t y0, int32_t x1, int32_t y1, int sh)
{
    const uint32_t tx0 = (uint32_t)x0 >> sh, tx1 = (uint32_t)(x1 - 1) >> sh;
    const uint32_t ty0 = (uint32_t)y0 >> sh, ty1 = (uint32_t)(y1 - 1) >> sh;
    return (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
}

// 32-bit two's-complement sum, the int32 arithmetic of the reference's
// point bounds as its compiled code performs it.
__device__ __forceinline__ int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

// ProcessPoints (CommandBuffer.Draw.cpp:1315-1383): centre
// (int)((p.x + 1) * 0.5f * (viewport.width - 1)), half extent
// (int)ceil(pointSize / 2), square [max(0, c - h), min(width, c + h + 1)).
__device__ void setup_point(const SetupArgs& a, const DrawDesc& d, uint32_t di, uint32_t p, uint32_t gp)
{
    uint32_t slot;
    if (!vertex_slot(d, p, &slot)) { a.counts[gp] = 0; return; }
    float P[4];
    clip_to_ndc(a.vout, a.stride, slot, P);
    const float size = d.psize >= 0 ? a.vout[(uint64_t)slot * a.stride + 4 + (uint32_t)d.psize] : 0.0f;
    const int32_t cx = trunc_to_i32((P[0] + 1.0f) * 0.5f * (d.vp_w - 1.0f));
    const int32_t cy = trunc_to_i32((P[1] + 1.0f) * 0.5f * (d.vp_h - 1.0f));
    const int32_t h = trunc_to_i32(ceilf(size / 2.0f));
    const int32_t x0 = max(0, wrap_sub(cx, h)), y0 = max(0, wrap_sub(cy, h));
    const int32_t x1 = min(d.limit_x, wrap_add(wrap_add(cx, h), 1)), y1 = min(d.limit_y, wrap_add(wrap_add(cy, h), 1));
    if (x0 >= x1 || y0 >= y1) { a.counts[gp] = 0; return; }
    TriSetup r;
    r.ax[0] = __int_as_float(cx);
    r.ay[0] = __int_as_float(cy);
    r.dx[0] = size;
    r.ax[1] = r.ax[2] = r.ay[1] = r.ay[2] = r.dx[1] = r.dx[2] = 0.0f;
    r.dy[0] = r.dy[1] = r.dy[2] = 0.0f;
    r.x0 = x0; r.y0 = y0; r.x1 = x1; r.y1 = y1;
    r.dt_lo = -__builtin_inff();
    r.dt_hi = __builtin_inff();
    r.area = 1.0f;
    r.draw = di;
    r.z[0] = P[2]; r.z[1] = 0.0f; r.z[2] = 0.0f;
    r.w[0] = r.w[1] = r.w[2] = P[3];
    r.v[0] = r.v[1] = r.v[2] = slot;
    r.flags = kPrimFront | kPrimPoint;   // DrawPixel(..., front = true, ...)
    r.provoking = slot;
    r.pad = 0;
    store_record(&a.recs[gp], r);
    a.counts[gp] = tiles_of(x0, y0, x1, y1, a.tile_shift);
}

// Pixel box that holds every pixel whose centre can pass the four edge
// tests of a line quad with corners c[4] (NDC), or false when the quad is
// too degenerate for the bound below (the caller then scans the viewport,
// as the reference does for every line).
//
// The bound: in pixel space (x + 0.5 = (ndc + 1) * extent / 2) let the quad
// have sides of at least 0.5 px and an angle with sine >= 0.1, corners
// within 4 of the origin in NDC and viewport extents <= 32768.  A pixel
// centre 3.5 px or more outside the corners' box is then at least
// 3.5 * sin(angle / 2) >= 0.16 px outside some edge, i.e. >= 9.6e-6 NDC.
// The float evaluation (x - ax) * dy - (y - ay) * dx of that edge errs by
// at most 8 * 2^-24 * (|x - ax| + |y - ay|) * |d| <= 4.8e-6 * |d|, so the
// edge value stays negative and the pixel fails.
__device__ bool line_box(const DrawDesc& d, const float* c[4], int32_t* x0, int32_t* y0, int32_t* x1, int32_t* y1)
{
    if (!(d.vp_w <= 32768.0f) || !(d.vp_h <= 32768.0f)) return false;
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (!(fabsf(c[k][0]) <= 4.0f) || !(fabsf(c[k][1]) <= 4.0f)) return false;
    const double sx = 0.5 * (double)d.vp_w, sy = 0.5 * (double)d.vp_h;
    // c = {p00, p01, p10, p11}: long side p00 -> p10, short side p00 -> p01
    const double ux = ((double)c[2][0] - (double)c[0][0]) * sx, uy = ((double)c[2][1] - (double)c[0][1]) * sy;
    const double vx = ((double)c[1][0] - (double)c[0][0]) * sx, vy = ((double)c[1][1] - (double)c[0][1]) * sy;
    const double lu = sqrt(ux * ux + uy * uy), lv = sqrt(vx * vx + vy * vy);
    if (!(lu >= 0.5) || !(lv >= 0.5) || !(fabs(ux * vy - uy * vx) >= 0.1 * lu * lv)) return false;
    double mnx = 1e30, mxx = -1e30, mny = 1e30, mxy = -1e30;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double px = ((double)c[k][0] + 1.0) * sx, py = ((double)c[k][1] + 1.0) * sy;
        mnx = fmin(mnx, px); mxx = fmax(mxx, px);
        mny = fmin(mny, py); mxy = fmax(mxy, py);
    }
    *x0 = max(*x0, (int32_t)floor(mnx) - 4);
    *y0 = max(*y0, (int32_t)floor(mny) - 4);
    *x1 = min(*x1, (int32_t)floor(mxx) + 5);
    *y1 = min(*y1, (int32_t)floor(mxy) + 5);
    return true;
}

// ProcessLines (CommandBuffer.Draw.cpp:1385-1508): the quad around p0p1
// with half widths LineWidth / (width, height) along (dir.y, -dir.x),
// dir = normalize(p1 - p0) over all four components; a pixel is drawn when
// its centre has E(p00, p01) >= 0, E(p11, p10) >= 0, E(p10, p00) >= 0 and
// E(p01, p11) >= 0.  The reference scans the whole viewport; the record's
// box is line_box's when it applies.
__device__ void setup_line(const SetupArgs& a, const DrawDesc& d, uint32_t di, uint32_t p, uint32_t gp)
{
    const uint32_t r0 = d.topology == 1 ? 2 * p : p;   // LINE_LIST or LINE_STRIP (CalculatePrimitives :584-611)
    uint32_t s0, s1;
    if (!vertex_slot(d, r0, &s0) || !vertex_slot(d, r0 + 1, &s1)) { a.counts[gp] = 0; return; }
    float P0[4], P1[4];
    clip_to_ndc(a.vout, a.stride, s0, P0);
    clip_to_ndc(a.vout, a.stride, s1, P1);
    const float lwx = d.line_width / d.vp_w, lwy = d.line_width / d.vp_h;
    const float vx = P1[0] - P0[0], vy = P1[1] - P0[1], vz = P1[2] - P0[2], vw = P1[3] - P0[3];
    const float inv = 1.0f / sqrtf(((vx * vx + vy * vy) + vz * vz) + vw * vw);   // glm::normalize (vec4)
    const float px = vy * inv, py = -(vx * inv);
    const float p00[2] = {P0[0] + px * lwx, P0[1] + py * lwy}, p01[2] = {P0[0] - px * lwx, P0[1] - py * lwy};
    const float p10[2] = {P1[0] + px * lwx, P1[1] + py * lwy}, p11[2] = {P1[0] - px * lwx, P1[1] - py * lwy};
    TriSetup r;
    put_edge(r, 0, p00, p01);
    put_edge(r, 1, p11, p10);
    put_edge(r, 2, p10, p00);
    LineExt L;
    L.ax3 = p01[0];
    L.ay3 = p01[1];
    L.dy3 = p11[1] - p01[1];
    L.dx3 = p11[0] - p01[0];
    L.p0x = P0[0];
    L.p0y = P0[1];
    L.ddx = vx;
    L.ddy = vy;
    const float len = sqrtf(vx * vx + vy * vy);   // glm::length(p1.xy - p0.xy)
    r.area = len * len;
    // ProcessLines loops while (float)x < viewport.width: ceil(width) columns
    const TargetParams& t = a.targets[d.target];
    int32_t x0 = 0, y0 = 0;
    int32_t x1 = d.vp_w >= (float)t.width ? (int32_t)t.width : (int32_t)ceilf(d.vp_w);
    int32_t y1 = d.vp_h >= (float)t.height ? (int32_t)t.height : (int32_t)ceilf(d.vp_h);
    const float* cs[4] = {p00, p01, p10, p11};
    line_box(d, cs, &x0, &y0, &x1, &y1);
    if (x0 >= x1 || y0 >= y1) { a.counts[gp] = 0; return; }
    r.x0 = x0; r.y0 = y0; r.x1 = x1; r.y1 = y1;
    r.dt_lo = -__builtin_inff();
    r.dt_hi = __builtin_inff();
    r.draw = di;
    r.z[0] = P0[2]; r.z[1] = P1[2]; r.z[2] = 0.0f;
    r.w[0] = P0[3]; r.w[1] = P1[3]; r.w[2] = P1[3];
    r.v[0] = s0; r.v[1] = s1; r.v[2] = s1;
    r.flags = kPrimFront | kPrimLine;
    r.provoking = s0;
    r.pad = 0;
    store_record(&a.recs[gp], r);
    a.ext[gp] = L;
    a.counts[gp] = tiles_of(x0, y0, x1, y1, a.tile_shift);
}

// Per-primitive part of ProcessTriangles (CommandBuffer.Draw.cpp:1526-1569)
// and the per-triangle constants of GetFragmentInput (:879-898, :900 cull).
__global__ __launch_bounds__(256) void setup_kernel(SetupArgs a)
{
    const uint32_t di = a.block_draw[blockIdx.x];
    const DrawDesc& d = a.draws[di];
    const uint32_t p = (blockIdx.x - d.sblock_base) * 256u + threadIdx.x;
    if (p >= d.prim_count) return;
    const uint32_t gp = d.prim_base + p;
    if (d.topology == 0) { setup_point(a, d, di, p, gp); return; }
    if (d.topology <= 2) { setup_line(a, d, di, p, gp); return; }

    // CalculatePrimitives (:567-673)
    uint32_t raw[3], prov;
    if (d.topology == 3) { raw[0] = 3 * p; raw[1] = 3 * p + 1; raw[2] = 3 * p + 2; prov = 3 * p; }
    else if (d.topology == 4) { raw[0] = p; raw[1] = p + 1; raw[2] = p + 2; prov = p; }
    else { raw[0] = 0; raw[1] = p + 1; raw[2] = p + 2; prov = p + 1; }

    uint32_t slot[4];
    const uint32_t rs[4] = {raw[0], raw[1], raw[2], prov};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int64_t s = rs[k];
        if (d.vmode == kSlotVertexId) {
            uint32_t idx;
            if (d.index_size == 1) idx = ((const uint8_t*)d.index_addr)[d.first + rs[k]];
            else if (d.index_size == 2) idx = ((const uint16_t*)d.index_addr)[d.first + rs[k]];
            else idx = ((const uint32_t*)d.index_addr)[d.first + rs[k]];
            s = (int64_t)d.vertex_offset + (int64_t)idx;
        }
        ok = ok && s >= 0 && s < (int64_t)d.vout_count;
        slot[k] = d.vout_base + (uint32_t)s;
    }
    if (!ok) { a.counts[gp] = 0; return; }   // vertex outside the bound buffer

    uint32_t i0 = slot[0], i1 = slot[1], i2 = slot[2];
    if (d.front_face == 1) { const uint32_t t = i0; i0 = i2; i2 = t; }   // CLOCKWISE swaps p0, p2

    float P[3][4];
    const uint32_t ids[3] = {i0, i1, i2};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float4 q = *(const float4*)(a.vout + (uint64_t)ids[k] * a.stride);
        P[k][0] = q.x / q.w;
        P[k][1] = q.y / q.w;
        P[k][2] = q.z / q.w;
        P[k][3] = q.w;
    }
    int32_t sx[3], sy[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        sx[k] = trunc_to_i32((P[k][0] + 1.0f) * 0.5f * d.vp_w);
        sy[k] = trunc_to_i32((P[k][1] + 1.0f) * 0.5f * d.vp_h);
    }
    const int32_t mnx = min(sx[0], min(sx[1], sx[2])), mxx = max(sx[0], max(sx[1], sx[2]));
    const int32_t mny = min(sy[0], min(sy[1], sy[2])), mxy = max(sy[0], max(sy[1], sy[2]));
    const int32_t x0 = max(0, mnx), y0 = max(0, mny);
    const int32_t x1 = (int32_t)min((int64_t)d.limit_x, (int64_t)mxx + 1);
    const int32_t y1 = (int32_t)min((int64_t)d.limit_y, (int64_t)mxy + 1);

    TriSetup r;
    float area = (P[2][0] - P[0][0]) * (P[1][1] - P[0][1]) - (P[2][1] - P[0][1]) * (P[1][0] - P[0][0]);
    bool front;
    if (area < 0) {
        area = -area;
        front = false;
        put_edge(r, 0, P[2], P[1]);
        put_edge(r, 1, P[0], P[2]);
        put_edge(r, 2, P[1], P[0]);
    } else {
        front = true;
        put_edge(r, 0, P[1], P[2]);
        put_edge(r, 1, P[2], P[0]);
        put_edge(r, 2, P[0], P[1]);
    }
    const bool culled = ((d.cull_mode & 2u) && !front) || ((d.cull_mode & 1u) && front);
    if (culled || x0 >= x1 || y0 >= y1) { a.counts[gp] = 0; return; }

    r.area = area;
    r.z[0] = P[0][2]; r.z[1] = P[1][2]; r.z[2] = P[2][2];
    r.w[0] = P[0][3]; r.w[1] = P[1][3]; r.w[2] = P[2][3];
    r.x0 = x0; r.y0 = y0; r.x1 = x1; r.y1 = y1;
    r.v[0] = i0; r.v[1] = i1; r.v[2] = i2;
    r.flags = front ? 1u : 0u;
    r.provoking = slot[3];
    r.draw = di;
    depth_bounds(r, P, d);
    uint4* dst = (uint4*)&a.recs[gp];
    const uint4* src = (const uint4*)&r;
#pragma unroll
    for (int q = 0; q < 8; q++) dst[q] = src[q];

    const int sh = a.tile_shift;
    const uint32_t tx0 = (uint32_t)x0 >> sh, tx1 = (uint32_t)(x1 - 1) >> sh;
    const uint32_t ty0 = (uint32_t)y0 >> sh, ty1 = (uint32_t)(y1 - 1) >> sh;
    a.counts[gp] = (tx1 - tx0 + 1) * (ty1 - ty0 + 1);
}

// ---------------------------------------------------------------- scans

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint32_t wave_base = 0;
    for (int w = 0; w < wave; w++) wave_base += s_tmp[w];
    total = s_tmp[0] + s_tmp[1] + s_tmp[2] + s_tmp[3];
    __syncthreads();
    return wave_base + x - v;
}

__device__ __forceinline__ uint64_t block_exclusive_scan64(uint64_t v, uint64_t* s_tmp, uint64_t& total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wave] = x;
    __syncthreads();
    uint64_t wave_base = 0;
    for (int w = 0; w < wave; w++) wave_base += s_tmp[w];
    total = s_tmp[0] + s_tmp[1] + s_tmp[2] + s_tmp[3];
    __syncthreads();
    return wave_base + x - v;
}

// Sum of each kScanTile-element slab, in 64 bits (a slab of large
// primitives' tile counts can pass 2^32).
__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* in, uint32_t n, uint64_t* sums)
{
    __shared__ uint64_t s_tmp[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) s += (base + i < n) ? in[base + i] : 0u;
    uint64_t total;
    block_exclusive_scan64(s, s_tmp, total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the slab sums in place by one workgroup; the grand
// total, in 64 bits, goes to *total64 (when given).
__global__ __launch_bounds__(kScanThreads) void scan_sums_kernel(uint64_t* sums, uint32_t n, uint64_t* total64)
{
    __shared__ uint64_t s_tmp[4];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < n; base += kScanThreads) {
        const uint32_t i = base + threadIdx.x;
        const uint64_t v = i < n ? sums[i] : 0u;
        uint64_t total;
        const uint64_t ex = block_exclusive_scan64(v, s_tmp, total);
        if (i < n) sums[i] = carry + ex;
        carry += total;
    }
    if (total64 && threadIdx.x == 0) *total64 = carry;
}

// Exclusive scan of every slab plus its slab offset; out[n] = grand total
// when write_total is set.  The 32-bit outputs are exact when the grand
// total is below 2^32 (the pair scan's is checked against that before its
// offsets are used).
__global__ __launch_bounds__(kScanThreads) void scan_down_kernel(const uint32_t* in, uint32_t n, const uint64_t* sums,
                                                                 uint32_t* out, int write_total)
{
    __shared__ uint32_t s_tmp[4];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? in[base + i] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = (uint32_t)sums[blockIdx.x] + block_exclusive_scan(s, s_tmp, total);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (write_total && base < n && base + kScanItems >= n) out[n] = run;
    if (write_total && n == 0 && blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
}

// ---------------------------------------------------------------- emit

struct EmitArgs {
    const TriSetup* recs;
    const uint32_t* counts;
    const uint32_t* offsets;
    const DrawDesc* draws;
    const TargetParams* targets;
    uint32_t* keys;               // global tile index per pair
    uint32_t* vals;               // primitive index per pair
    uint32_t* big;                // primitives left to emit_big_kernel; big[0] is the count
    uint32_t prims;
    int32_t tile_shift;
    uint32_t cap;                 // capacity of keys / vals; pairs at or beyond it are dropped
};

constexpr uint32_t kEmitBig = 2048;   // pairs above which a primitive goes to emit_big_kernel

// Block bits of a pair value (kPairBlockShift): which quarters of tile
// (col, row) of the primitive's tile rectangle its pixel box overlaps (the
// 8x8 blocks of a 16x16 tile, the 16x16 quadrants of a 32x32 tile); box is
// relative to the rectangle's first tile.  0 for 8x8 tiles.
__device__ __forceinline__ uint32_t pair_blocks(int tile_shift, int4 box, uint32_t col, uint32_t row)
{
    if (tile_shift != 4 && tile_shift != 5) return 0u;
    const int half = 1 << (tile_shift - 1);
    const int X = (int)col << tile_shift, Y = (int)row << tile_shift;
    uint32_t m = 0u;
#pragma unroll
    for (int by = 0; by < 2; by++)
#pragma unroll
        for (int bx = 0; bx < 2; bx++) {
            const int x0 = X + bx * half, y0 = Y + by * half;
            if (box.x < x0 + half && box.z > x0 && box.y < y0 + half && box.w > y0) m |= 1u << (by * 2 + bx);
        }
    return m << kPairBlockShift;
}

// Pixel centre in NDC as the raster kernel computes it (ProcessTriangles,
// CommandBuffer.Draw.cpp:1571-1580).
__device__ __forceinline__ float pixel_centre(int p, float extent, float half)
{
    return ((float)p / extent + half) * 2.0f - 1.0f;
}

// An interval [lo, hi] holding pixel_centre(p) for every p0 <= p <= p1 (0
// <= p < 2^24), from multiplications by 2 * half = fl(1 / extent) instead
// of divisions.  fl(p * fl(1/e)) is within 1.5 * 2^-23 |p / e| of
// fl(p / e), and the rounded steps after it (+ half, * 2, - 1) keep the
// difference within 2^-23 (3.5 |c| + 7 |half| + 3.5) of the exact centre
// c, which the margin 2^-20 (|c'| + 4) exceeds, with room for its own and
// the subtraction's rounding, when |half| <= 1, i.e. |extent| >= 0.5 (the
// caller checks).  The centres are monotone in p, so every one lies
// between the two ends'.
__device__ __forceinline__ void centre_range(int p0, int p1, float half, float& lo, float& hi)
{
    const float inv = 2.0f * half;
    const float a = ((float)p0 * inv + half) * 2.0f - 1.0f;
    const float b = ((float)p1 * inv + half) * 2.0f - 1.0f;
    const float ea = fmaf(fabsf(a), 0x1p-20f, 0x1p-18f);
    const float eb = fmaf(fabsf(b), 0x1p-20f, 0x1p-18f);
    lo = fminf(a - ea, b - eb);
    hi = fmaxf(a + ea, b + eb);
}

// pair_blocks less, for a triangle, the quarters where it covers no pixel
// (points and lines keep their box's quarters).  box is relative to the
// primitive's first tile, whose pixel origin is (ox, oy); e holds the
// record's edges (bytes 0-47: ax[3], ay[3], dy[3], dx[3]) and vp the draw's
// (vp_w, vp_h, half_x, half_y).
//
// The raster kernel's edge value at a pixel centre (x, y) is
// fl(A(x) - B(y)), A(x) = fl(fl(x - a.x) * dy), B(y) = fl(fl(y - a.y) * dx),
// fp32 without contraction, at the centres pixel_centre gives.  Rounding is
// monotone, so with finite edge terms and centres A is non-decreasing in x
// when dy > 0, non-increasing when dy < 0 and zero when dy = 0, B likewise
// in y with dx, and fl(A - B) is non-decreasing in A and non-increasing in
// B.  Over a quarter's centres, x in [xlo, xhi] and y in [ylo, yhi], no
// pixel's value exceeds A(dy > 0 ? xhi : xlo) - B(dx < 0 ? yhi : ylo); when
// that is negative every pixel's is (a pixel's NaN needs A = B = +-inf,
// which makes the bound +inf or NaN) and the quarter is dropped.  NaN is
// not < 0, and non-finite terms keep every quarter.  Each edge's bound is
// computed once per quarter column and row, from an interval holding the
// column's (row's) centres: centre_range's, or for viewport extents below
// 0.5 in magnitude the exact centres of its two ends.
__device__ __forceinline__ uint32_t pair_blocks_covered(int tile_shift, int4 box, uint32_t col, uint32_t row, bool triangle,
                                                        const float4 e[3], float4 vp, int ox, int oy)
{
    const uint32_t m = pair_blocks(tile_shift, box, col, row);
    if (m == 0u || !triangle) return m;
    const int half = 1 << (tile_shift - 1);
    const int X = (int)col << tile_shift, Y = (int)row << tile_shift;
    const uint32_t bits = m >> kPairBlockShift;
    const float ax[3] = {e[0].x, e[0].y, e[0].z}, ay[3] = {e[0].w, e[1].x, e[1].y};
    const float dy[3] = {e[1].z, e[1].w, e[2].x}, dx[3] = {e[2].y, e[2].z, e[2].w};
    bool finite = true;   // & rather than &&: no branch per term
#pragma unroll
    for (int i = 0; i < 3; i++)
        finite = finite & __builtin_isfinite(ax[i]) & __builtin_isfinite(ay[i]) & __builtin_isfinite(dx[i]) &
                 __builtin_isfinite(dy[i]);
    // NaN extents fail the test and take the exact centres
    const bool approx = fabsf(vp.x) >= 0.5f && fabsf(vp.y) >= 0.5f;
    auto range = [&](int p0, int p1, float extent, float hf, float& lo, float& hi) {
        if (approx) {
            centre_range(p0, p1, hf, lo, hi);
        } else {
            const float u = pixel_centre(p0, extent, hf), v = pixel_centre(p1, extent, hf);
            lo = fminf(u, v);
            hi = fmaxf(u, v);
            finite = finite & __builtin_isfinite(u) & __builtin_isfinite(v);
        }
        finite = finite & __builtin_isfinite(lo) & __builtin_isfinite(hi);
    };
    float xlo[2], xhi[2], ylo[2], yhi[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        xlo[h] = xhi[h] = ylo[h] = yhi[h] = 0.0f;
        // quarter column h is in the box when a quarter of it is (bits h, h + 2)
        if (bits & (5u << h))
            range(ox + max(X + half * h, box.x), ox + min(X + half * h + half, box.z) - 1, vp.x, vp.z, xlo[h], xhi[h]);
        // quarter row h (bits 2h, 2h + 1)
        if (bits & (3u << (2 * h)))
            range(oy + max(Y + half * h, box.y), oy + min(Y + half * h + half, box.w) - 1, vp.y, vp.w, ylo[h], yhi[h]);
    }
    if (!finite) return m;
    float amax[3][2], bmin[3][2];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int h = 0; h < 2; h++) {
            amax[i][h] = ((dy[i] > 0.0f ? xhi[h] : xlo[h]) - ax[i]) * dy[i];
            bmin[i][h] = ((dx[i] < 0.0f ? yhi[h] : ylo[h]) - ay[i]) * dx[i];
        }
    uint32_t keep = 0u;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        bool cover = ((bits >> b) & 1u) != 0u;
#pragma unroll
        for (int i = 0; i < 3; i++) cover = cover & !(amax[i][b & 1] - bmin[i][b >> 1] < 0.0f);
        if (cover) keep |= 1u << (kPairBlockShift + b);
    }
    return keep;
}

// Writes the (tile, primitive) pairs of every primitive at its scanned
// offset, so the pair array is in primitive order.  A workgroup takes 256
// consecutive primitives, whose pairs are one contiguous run of the output
// (less the runs of primitives with more than kEmitBig pairs, queued for
// emit_big_kernel): its threads stride over the run, so consecutive lanes
// write consecutive pairs, and each finds its primitive by binary search of
// the workgroup's scanned counts in LDS.  What a pair needs of its
// primitive (tile rectangle, edges, viewport) is staged in LDS by the
// primitive's thread, so the pair loop reads no record.  Tiles of a
// primitive go in row-major order of its tile rectangle.
__global__ __launch_bounds__(256) void emit_kernel(EmitArgs a)
{
    constexpr uint32_t kNoEdgeTest = 0x80000000u;   // s_w flag: a point or line keeps its box's quarters
    __shared__ uint32_t s_tmp[4];
    __shared__ uint32_t s_off[257];   // exclusive scan of the kept pair counts
    __shared__ uint32_t s_dst[256];   // output offset of each primitive's first pair
    __shared__ uint32_t s_key[256];   // global tile index of each primitive's first tile
    __shared__ uint32_t s_w[256];     // tile columns of each primitive's rectangle, | kNoEdgeTest
    __shared__ uint32_t s_tx[256];    // tile columns of each primitive's target
    __shared__ int4 s_box[256];       // pixel box of each primitive, relative to its first tile
    __shared__ int2 s_org[256];       // pixel origin of each primitive's first tile
    __shared__ float4 s_edge[256][3]; // each primitive's edges (record bytes 0-47)
    __shared__ float4 s_vp[256];      // its draw's vp_w, vp_h, half_x, half_y
    const uint32_t nblk = (a.prims + 255u) / 256u;
    for (uint32_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const uint32_t gp = blk * 256u + threadIdx.x;
        uint32_t c = gp < a.prims ? a.counts[gp] : 0u;
        if (c > kEmitBig) {
            a.big[1 + atomicAdd(&a.big[0], 1u)] = gp;
            c = 0;
        }
        if (c) {
            const TriSetup& r = a.recs[gp];
            const float4* e4 = (const float4*)&r;
            const float4 e0 = e4[0], e1 = e4[1], e2 = e4[2];
            const DrawDesc& dd = a.draws[r.draw];
            const TargetParams& t = a.targets[dd.target];
            const int sh = a.tile_shift;
            const uint32_t tx0 = (uint32_t)r.x0 >> sh, tx1 = (uint32_t)(r.x1 - 1) >> sh;
            const uint32_t ty0 = (uint32_t)r.y0 >> sh;
            s_dst[threadIdx.x] = a.offsets[gp];
            s_key[threadIdx.x] = t.tile_base + ty0 * t.tiles_x + tx0;
            s_w[threadIdx.x] = (tx1 - tx0 + 1) | ((r.flags & (kPrimPoint | kPrimLine)) ? kNoEdgeTest : 0u);
            s_tx[threadIdx.x] = t.tiles_x;
            const int ox = (int)(tx0 << sh), oy = (int)(ty0 << sh);
            s_box[threadIdx.x] = make_int4(r.x0 - ox, r.y0 - oy, r.x1 - ox, r.y1 - oy);
            s_org[threadIdx.x] = make_int2(ox, oy);
            s_edge[threadIdx.x][0] = e0;
            s_edge[threadIdx.x][1] = e1;
            s_edge[threadIdx.x][2] = e2;
            s_vp[threadIdx.x] = make_float4(dd.vp_w, dd.vp_h, dd.half_x, dd.half_y);
        }
        uint32_t total;
        s_off[threadIdx.x] = block_exclusive_scan(c, s_tmp, total);
        if (threadIdx.x == 0) s_off[256] = total;
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < total; j += 256u) {
            // lo: the last primitive whose run starts at or before j (s_off is
            // non-decreasing and s_off[0] = 0), in eight unrolled halvings
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = 128; step >= 1; step >>= 1)
                if (s_off[lo + step] <= j) lo += step;
            const uint32_t k = j - s_off[lo];
            const uint32_t wl = s_w[lo];
            const uint32_t w = wl & ~kNoEdgeTest;
            // row = k / w: k < kEmitBig = 2048 and w <= 2048, so (k + 1/2) / w
            // is at least 1/(2w) from an integer while the product with
            // v_rcp_f32's reciprocal (1 ulp) errs by under 1.5 * 2^-23 * 2049 / w
            const uint32_t row = (uint32_t)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)w));
            const uint32_t out = s_dst[lo] + k;
            if (out < a.cap) {
                const uint32_t col = k - row * w;
                a.keys[out] = s_key[lo] + row * s_tx[lo] + col;
                const int2 org = s_org[lo];
                const float4 e[3] = {s_edge[lo][0], s_edge[lo][1], s_edge[lo][2]};
                a.vals[out] = (blk * 256u + lo) |
                              pair_blocks_covered(a.tile_shift, s_box[lo], col, row, (wl & kNoEdgeTest) == 0u, e, s_vp[lo], org.x, org.y);
            }
        }
        __syncthreads();
    }
}

// Pairs of the large primitives: one workgroup per primitive at a time,
// its threads striding over the primitive's tile rectangle (row-major, the
// order emit_kernel uses).
__global__ __launch_bounds__(256) void emit_big_kernel(EmitArgs a)
{
    const uint32_t n = a.big[0];
    for (uint32_t b = blockIdx.x; b < n; b += gridDim.x) {
        const uint32_t gp = a.big[1 + b];
        const TriSetup& r = a.recs[gp];
        const TargetParams& t = a.targets[a.draws[r.draw].target];
        const int sh = a.tile_shift;
        const uint32_t tx0 = (uint32_t)r.x0 >> sh, tx1 = (uint32_t)(r.x1 - 1) >> sh;
        const uint32_t ty0 = (uint32_t)r.y0 >> sh;
        const uint32_t w = tx1 - tx0 + 1;
        const uint32_t c = a.counts[gp], off = a.offsets[gp];
        const int4 box = make_int4(r.x0 - (int)(tx0 << sh), r.y0 - (int)(ty0 << sh), r.x1 - (int)(tx0 << sh),
                                   r.y1 - (int)(ty0 << sh));
        const DrawDesc& dd = a.draws[r.draw];
        const float4* e4 = (const float4*)&r;
        const float4 e[3] = {e4[0], e4[1], e4[2]};
        const float4 vp = make_float4(dd.vp_w, dd.vp_h, dd.half_x, dd.half_y);
        const bool triangle = (r.flags & (kPrimPoint | kPrimLine)) == 0u;
        for (uint32_t k = threadIdx.x; k < c && off + k < a.cap; k += 256u) {
            const uint32_t row = k / w, col = k % w;
            a.keys[off + k] = t.tile_base + (ty0 + row) * t.tiles_x + tx0 + col;
            a.vals[off + k] = gp | pair_blocks_covered(sh, box, col, row, triangle, e, vp, (int)(tx0 << sh), (int)(ty0 << sh));
        }
    }
}

// Pair count kept for sorting: the 64-bit scanned total clamped to the
// capacity of the pair arrays; a larger total raises the overflow word, and
// then the emit kernels write only the pairs below the capacity and the
// raster kernel draws nothing.
__global__ void clamp_pairs_kernel(const uint64_t* total, uint32_t cap, uint32_t* kept, uint32_t* overflow)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint64_t t = *total;
    *kept = t < (uint64_t)cap ? (uint32_t)t : cap;
    if (t > (uint64_t)cap) *overflow = 1u;
}

// First index in [lo, hi) of the ascending keys whose tile is not before
// t; hi when there is none.
__device__ __forceinline__ uint32_t lower_bound_key(const uint32_t* keys, uint32_t lo, uint32_t hi, uint32_t t)
{
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (keys[mid] < t) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Start of every tile's range in the tile-sorted pair array, start[t] =
// the first pair whose tile is not before t (start[tiles] = pair count).
// A workgroup takes 256 consecutive tiles: two searches over all pairs bound
// its tiles' ranges, and each thread searches only between them (empty
// tiles cost nothing, however many a sparse frame has).  The pair count is
// read on the device.
__global__ __launch_bounds__(256) void tile_ranges_kernel(const uint32_t* keys, const uint32_t* n_ptr, uint32_t tiles, uint32_t* start)
{
    __shared__ uint32_t s_lim[2];
    const uint32_t t0 = blockIdx.x * 256u;
    const uint32_t n = *n_ptr;
    if (threadIdx.x < 2) s_lim[threadIdx.x] = lower_bound_key(keys, 0u, n, min(t0 + threadIdx.x * 256u, tiles + 1u));
    __syncthreads();
    const uint32_t t = t0 + threadIdx.x;
    if (t <= tiles) start[t] = lower_bound_key(keys, s_lim[0], s_lim[1], t);
}

// ---------------------------------------------------------- radix sort

// A pass sorts by an 8-bit digit.  The pair array is cut into kSortSegments
// contiguous segments, one per workgroup; the histogram kernel counts each
// segment's digits, a scan over (digit, segment) turns the counts into
// output bases, and the scatter kernel walks its segment in order in tiles
// of kSortTile items.  Each wave ranks its quarter of a tile (64-item
// rounds, equal digits ranked with ballots), the workgroup orders the tile
// by digit in LDS -- stable: by wave, then round, then lane, which is the
// items' order -- and the sorted tile is written out with consecutive
// threads on consecutive positions of each digit's run, so the stores of a
// run coalesce instead of scattering one item per digit bucket.
constexpr int kSortSegments = 1024;
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortRounds = 8;                              // 64-item rounds per wave and tile
constexpr int kSortTile = kSortThreads * kSortRounds;       // 2048 items

__device__ __forceinline__ void segment_range(uint32_t n, uint32_t seg, uint32_t& b, uint32_t& e)
{
    const uint32_t per = ((n + kSortSegments - 1) / kSortSegments + 63u) & ~63u;
    b = min(n, seg * per);
    e = min(n, b + per);
}

__global__ __launch_bounds__(kSortThreads) void radix_hist_kernel(const uint32_t* keys, const uint32_t* n_ptr, int shift,
                                                                 uint32_t* hist)
{
    __shared__ uint32_t s_cnt[kSortWaves][256];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) s_cnt[w][threadIdx.x] = 0;
    __syncthreads();
    uint32_t b, e;
    segment_range(*n_ptr, blockIdx.x, b, e);
    // kSortRounds keys per thread loaded before they are counted, so the loads overlap
    for (uint32_t i0 = b; i0 < e; i0 += (uint32_t)kSortTile) {
        uint32_t ks[kSortRounds];
#pragma unroll
        for (int r = 0; r < kSortRounds; r++) {
            const uint32_t i = i0 + (uint32_t)r * kSortThreads + threadIdx.x;
            ks[r] = i < e ? keys[i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < kSortRounds; r++)
            if (i0 + (uint32_t)r * kSortThreads + threadIdx.x < e) atomicAdd(&s_cnt[wave][(ks[r] >> shift) & 255u], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; w++) c += s_cnt[w][threadIdx.x];
    hist[threadIdx.x * kSortSegments + blockIdx.x] = c;
}

__global__ __launch_bounds__(kSortThreads) void radix_scatter_kernel(const uint32_t* keys_in, const uint32_t* vals_in,
                                                                    uint32_t* keys_out, uint32_t* vals_out,
                                                                    const uint32_t* n_ptr, int shift, const uint32_t* base)
{
    __shared__ uint32_t s_cnt[kSortWaves][256];   // a wave's items per digit, then the wave's offset in the digit's run
    __shared__ uint32_t s_start[256];             // where a digit's run starts in the sorted tile
    __shared__ uint32_t s_run[256];               // the segment's next output position per digit
    __shared__ uint32_t s_keys[kSortTile], s_vals[kSortTile];
    __shared__ uint64_t s_tmp[kSortWaves];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t d_me = threadIdx.x;            // the digit this thread keeps the books of
    s_run[d_me] = base[d_me * kSortSegments + blockIdx.x];
    uint32_t b, e;
    const uint32_t n_total = *n_ptr;
    segment_range(n_total, blockIdx.x, b, e);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (uint32_t t0 = b; t0 < e; t0 += (uint32_t)kSortTile) {
#pragma unroll
        for (int w = 0; w < kSortWaves; w++) s_cnt[w][d_me] = 0;
        __syncthreads();
        // this wave's quarter of the tile: kSortRounds rounds of 64 consecutive items
        const uint32_t w0 = t0 + (uint32_t)wave * (64u * kSortRounds);
        uint32_t ks[kSortRounds], vs[kSortRounds], rk[kSortRounds];
#pragma unroll
        for (int r = 0; r < kSortRounds; r++) {
            const uint32_t i = w0 + (uint32_t)r * 64u + (uint32_t)lane;
            ks[r] = i < e ? keys_in[i] : 0u;
            vs[r] = i < e ? vals_in[i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < kSortRounds; r++) {
            const bool valid = w0 + (uint32_t)r * 64u + (uint32_t)lane < e;
            const uint32_t dg = (ks[r] >> shift) & 255u;
            // lanes of this round holding the same digit
            uint64_t match = __ballot(valid);
#pragma unroll
            for (int bit = 0; bit < 8; bit++) {
                const bool set = (dg >> bit) & 1u;
                const uint64_t bal = __ballot(valid && set);
                match &= set ? bal : ~bal;
            }
            const int leader = __ffsll((long long)match) - 1;
            uint32_t run = 0;
            if (valid) run = s_cnt[wave][dg];
            __builtin_amdgcn_wave_barrier();
            if (valid && lane == leader) s_cnt[wave][dg] = run + (uint32_t)__popcll(match);
            __builtin_amdgcn_wave_barrier();
            rk[r] = run + (uint32_t)__popcll(match & lt);
        }
        __syncthreads();
        // per digit: the waves' offsets within the digit's run, and its total
        uint32_t total = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; w++) {
            const uint32_t c = s_cnt[w][d_me];
            s_cnt[w][d_me] = total;
            total += c;
        }
        uint64_t tile_n;
        const uint32_t start = (uint32_t)block_exclusive_scan64(total, s_tmp, tile_n);
        s_start[d_me] = start;
        __syncthreads();
        // the tile in digit order, in LDS
#pragma unroll
        for (int r = 0; r < kSortRounds; r++) {
            if (w0 + (uint32_t)r * 64u + (uint32_t)lane < e) {
                const uint32_t dg = (ks[r] >> shift) & 255u;
                const uint32_t pos = s_start[dg] + s_cnt[wave][dg] + rk[r];
                s_keys[pos] = ks[r];
                s_vals[pos] = vs[r];
            }
        }
        __syncthreads();
        // consecutive threads write consecutive positions of a digit's run
#pragma unroll
        for (int r = 0; r < kSortRounds; r++) {
            const uint32_t p = (uint32_t)r * kSortThreads + threadIdx.x;
            if (p < (uint32_t)tile_n) {
                const uint32_t key = s_keys[p];
                const uint32_t dg = (key >> shift) & 255u;
                const uint32_t dst = s_run[dg] + (p - s_start[dg]);
                if (dst < n_total) {
                    keys_out[dst] = key;
                    vals_out[dst] = s_vals[p];
                }
            }
        }
        __syncthreads();
        s_run[d_me] += total;
    }
}

}  // namespace cpv
