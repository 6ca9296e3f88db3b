// Host runtime of libcpv.so, part 2: the launch sequence.
//
// A render call becomes, per run of draws that share a shader program, one
// launch of each stage on the context's stream:
//   vertex kernel -> setup kernel -> scan -> emit -> radix sort by tile ->
//   tile ranges -> [tile features -> cost prediction -> queue order] ->
//   raster kernel
// The raster kernel fuses the render pass clear (loadOp CLEAR) into its tile
// initialisation, so a frame touches each framebuffer byte once on the way
// in (LOAD only) and once on the way out.
//
// The one device value the host needs -- the number of (tile, primitive)
// pairs, which sizes the sort -- is never waited for before the raster
// kernel: the pair arrays have a capacity taken from earlier launches (or,
// in a recorded batch, from the recording), the emit kernels drop pairs past
// it and raise an overflow word instead of writing out of bounds, the raster
// kernels that see the word draw nothing, and the next call on the context
// checks the count and redraws what did not fit.  Everything before the
// raster kernel (the front half) runs on the context's second stream, so a
// launch's front half overlaps the previous launch's raster kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <thread>
#include <cstring>
#include <mutex>
#include <vector>

#include "cpv_internal.h"
#include "raster_kernels.h"
#include "bin_kernels.h"
#include "sched_kernels.h"

using namespace cpv;
using namespace cpvrt;

namespace {

// Host staging of one or more launches' descriptors, copied to the device
// arena in one transfer.
struct Staging {
    std::vector<uint8_t> bytes;
    size_t put(const void* p, size_t n, size_t align = 256)
    {
        size_t at = (bytes.size() + align - 1) & ~(align - 1);
        bytes.resize(at + n);
        if (n) std::memcpy(bytes.data() + at, p, n);
        return at;
    }
};

struct LaunchDraw {
    const cpv_pipeline_desc* pipe;
    const cpv_draw_desc* draw;
    uint32_t instance;
};

// Everything one launch needs, resolved on the host.
struct LaunchPlan {
    Program prog;
    std::vector<DrawDesc> dd;
    std::vector<DrawEnv> envs;
    std::vector<FragmentState> fstates;
    std::vector<TargetParams> targets;
    std::vector<uint32_t> vblock, sblock;
    uint32_t vout_total = 0, prim_total = 0, tiles = 0, stride = 0;
    int tile_shift = 4;
    int raster_variant = 3;          // index into Program::raster
    bool lines = false;              // some draw rasterises lines (LineExt records)
    int32_t hiz_op = -1, hiz_kind = 32;
    int32_t uniform_fstate = -1;     // every draw's fragment state is fstates[0] (0), or not (-1)
    size_t o_env = 0, o_dd = 0, o_fs = 0, o_tp = 0, o_vb = 0, o_sb = 0;   // offsets in the arena
};

int plan_launch(const cpv_config& cfg, const Program& prog, const std::vector<TargetParams>& targets,
                const std::vector<LaunchDraw>& lds, const std::vector<cpv_pass_desc>& passes,
                const std::vector<uint32_t>& draw_target, LaunchPlan* P)
{
    const int tile = (int)cfg.tile_size;
    P->prog = prog;
    P->targets = targets;
    P->tile_shift = tile == 8 ? 3 : tile == 16 ? 4 : 5;
    P->stride = 4u + (uint32_t)prog.vs_outputs;
    if (prog.fs_inputs > prog.vs_outputs) return fail(CPV_ERROR_INVALID, "fragment inputs exceed vertex outputs");
    P->dd.assign(lds.size(), DrawDesc{});
    P->envs.resize(lds.size());
    P->fstates.resize(lds.size());
    uint32_t vout_total = 0, prim_total = 0;
    for (size_t i = 0; i < lds.size(); i++) {
        const cpv_pipeline_desc& p = *lds[i].pipe;
        const cpv_draw_desc& d = *lds[i].draw;
        int rc = build_env(p, d, &P->envs[i]);
        if (rc) return rc;
        rc = build_fragment_state(p, passes[draw_target[i]], &P->fstates[i]);
        if (rc) return rc;
        if (p.topology > CPV_TOPOLOGY_TRIANGLE_FAN)
            return fail(CPV_ERROR_UNSUPPORTED, "topology %u (adjacency and patch lists)", p.topology);
        if (p.topology < CPV_TOPOLOGY_TRIANGLE_LIST && (cfg.tile_size != 16 || cfg.waves_per_tile != 4))
            return fail(CPV_ERROR_INVALID, "points and lines raster with 16x16 tiles and 4 waves");
        if (p.topology != CPV_TOPOLOGY_POINT_LIST && p.topology < CPV_TOPOLOGY_TRIANGLE_LIST) P->lines = true;
        if (d.indexed && d.index_size != 1 && d.index_size != 2 && d.index_size != 4)
            return fail(CPV_ERROR_INVALID, "index size %u", d.index_size);
        if (d.indexed && p.primitive_restart) return fail(CPV_ERROR_UNSUPPORTED, "primitive restart");
        if (!(d.viewport[2] > 0.0f) || !(d.viewport[3] > 0.0f)) return fail(CPV_ERROR_INVALID, "viewport extent");
        if (!p.program) {
            // resources the built-in shaders dereference
            if ((p.vs_kind == CPV_VS_MVP_COLOR || p.vs_kind == CPV_VS_MVP_LAMBERT) && !d.uniform_address[0])
                return fail(CPV_ERROR_INVALID, "vertex shader needs a uniform buffer at binding 0");
            if (p.fs_kind == CPV_FS_TEX_LAMBERT && (!d.uniform_address[2] || !d.textures[1].address))
                return fail(CPV_ERROR_INVALID, "fragment shader needs a uniform at binding 2 and a texture at binding 1");
        }
        DrawDesc& o = P->dd[i];
        std::memset(&o, 0, sizeof(o));
        o.target = draw_target[i];
        o.fstate = (uint32_t)i;
        o.topology = p.topology;
        o.front_face = p.front_face;
        o.cull_mode = p.cull_mode;
        o.first = d.first;
        o.vertex_offset = d.vertex_offset;
        o.index_size = d.index_size;
        o.index_addr = d.index_address;
        o.instance_id = d.first_instance + lds[i].instance;
        o.vp_w = d.viewport[2];
        o.vp_h = d.viewport[3];
        o.vp_min_depth = d.viewport[4];
        o.vp_max_depth = d.viewport[5];
        o.half_x = 1.0f / d.viewport[2] * 0.5f;
        o.half_y = 1.0f / d.viewport[3] * 0.5f;
        o.line_width = p.line_width;
        o.psize = prog.psize;
        const float vw = d.viewport[2], vh = d.viewport[3];
        const int32_t lx = (vw >= -2147483648.0f && vw < 2147483648.0f) ? (int32_t)vw : INT32_MIN;
        const int32_t ly = (vh >= -2147483648.0f && vh < 2147483648.0f) ? (int32_t)vh : INT32_MIN;
        const TargetParams& t = targets[o.target];
        o.limit_x = std::min<int32_t>(lx, (int32_t)t.width);
        o.limit_y = std::min<int32_t>(ly, (int32_t)t.height);
        // vertex slots
        uint32_t nverts = d.count;
        o.vmode = d.indexed ? kSlotPerIndex : kSlotSequential;
        if (d.indexed) {
            // Shade every vertex of the bound buffers once when that is fewer
            // invocations than one per index (results are identical: the
            // vertex shader depends only on the vertex and instance ids).
            uint64_t nv = UINT64_MAX;
            bool any = false;
            for (uint32_t a = 0; a < p.attribute_count; a++) {
                const cpv_vertex_attribute_desc& at = p.attributes[a];
                for (uint32_t b = 0; b < p.binding_count; b++) {
                    if (p.bindings[b].binding != at.binding || p.bindings[b].input_rate != CPV_INPUT_RATE_VERTEX) continue;
                    if (p.bindings[b].stride == 0) continue;
                    any = true;
                    const uint64_t bytes = d.vertex_bytes[at.binding & 15];
                    nv = std::min<uint64_t>(nv, bytes / p.bindings[b].stride);
                }
            }
            if (any && nv < d.count) { o.vmode = kSlotVertexId; nverts = (uint32_t)nv; }
        }
        o.vout_base = vout_total;
        o.vout_count = nverts;
        o.prim_base = prim_total;
        o.prim_count = prim_count_of(p.topology, d.count);
        o.vblock_base = (uint32_t)P->vblock.size();
        o.sblock_base = (uint32_t)P->sblock.size();
        for (uint32_t b = 0; b < (nverts + 255) / 256; b++) P->vblock.push_back((uint32_t)i);
        for (uint32_t b = 0; b < (o.prim_count + 255) / 256; b++) P->sblock.push_back((uint32_t)i);
        vout_total += nverts;
        prim_total += o.prim_count;
    }
    if (prim_total >= (1u << 28)) return fail(CPV_ERROR_UNSUPPORTED, "more than 2^28 primitives in one launch");
    P->vout_total = vout_total;
    P->prim_total = prim_total;
    bool any_prim = false;
    for (const DrawDesc& o : P->dd) any_prim = any_prim || o.topology < CPV_TOPOLOGY_TRIANGLE_LIST;
    P->raster_variant = any_prim ? kRasterAnyPrim : tile_config_index(cfg.tile_size, cfg.waves_per_tile);
    // the fixed-format variant when folding the formats changes no pipeline
    // or target of the launch
    if (!any_prim) {
        bool fixed = true;
        for (const FragmentState& f : P->fstates) {
            FragmentState g = f;
            fold_fixed_formats(g);
            fixed = fixed && std::memcmp(&g, &f, sizeof(f)) == 0;
        }
        for (const TargetParams& t : P->targets) {
            TargetParams g = t;
            fold_fixed_formats(g);
            fixed = fixed && std::memcmp(&g, &t, sizeof(t)) == 0;
        }
        if (fixed) P->raster_variant = kRasterFixedFormats + tile_config_index(cfg.tile_size, cfg.waves_per_tile);
    }
    uint32_t tiles = 0;
    for (const TargetParams& t : targets) tiles = std::max(tiles, t.tile_base + t.tiles_x * t.tiles_y);
    P->tiles = tiles;
    // Hierarchical depth culling is exact when only the depth test observes
    // a fragment (no stencil, no depth bounds), every draw uses the same
    // compare op and depth format, and nobody counts covered fragments.
    P->hiz_op = -1;
    P->hiz_kind = 32;
    if (!cfg.count_fragments && !P->fstates.empty()) {
        const std::vector<FragmentState>& fs = P->fstates;
        bool ok = true;
        for (const FragmentState& f : fs) {
            ok = ok && f.depth_test && f.has_ds && f.ds.depth_off >= 0 && f.ds.vk_format != CPV_FORMAT_S8_UINT &&
                 !f.depth_bounds_test && !f.stencil_test && f.depth_compare == fs[0].depth_compare &&
                 f.ds.depth_kind == fs[0].ds.depth_kind;
        }
        for (const TargetParams& t : targets) ok = ok && t.has_ds && t.ds_fmt.depth_kind == fs[0].ds.depth_kind;
        if (ok) {
            P->hiz_op = fs[0].depth_compare;
            P->hiz_kind = fs[0].ds.depth_kind;
        }
    }
    P->uniform_fstate = P->fstates.empty() ? -1 : 0;
    for (const FragmentState& f : P->fstates)
        if (std::memcmp(&f, &P->fstates[0], sizeof(f)) != 0) P->uniform_fstate = -1;
    // the block tables are never empty
    if (P->vblock.empty()) P->vblock.push_back(0);
    if (P->sblock.empty()) P->sblock.push_back(0);
    return CPV_OK;
}

// Places a plan's descriptors in the staging buffer (draw descriptors are
// written again by patch_plan once the arena address is known).
void stage_plan(LaunchPlan& P, Staging& st)
{
    P.o_env = st.put(P.envs.data(), P.envs.size() * sizeof(DrawEnv));
    P.o_dd = st.put(P.dd.data(), P.dd.size() * sizeof(DrawDesc));
    P.o_fs = st.put(P.fstates.data(), P.fstates.size() * sizeof(FragmentState));
    P.o_tp = st.put(P.targets.data(), P.targets.size() * sizeof(TargetParams));
    P.o_vb = st.put(P.vblock.data(), P.vblock.size() * 4);
    P.o_sb = st.put(P.sblock.data(), P.sblock.size() * 4);
}

void patch_plan(LaunchPlan& P, Staging& st, uint8_t* arena)
{
    for (size_t i = 0; i < P.dd.size(); i++) P.dd[i].env = (uint64_t)(uintptr_t)(arena + P.o_env + i * sizeof(DrawEnv));
    if (!P.dd.empty()) std::memcpy(st.bytes.data() + P.o_dd, P.dd.data(), P.dd.size() * sizeof(DrawDesc));
}

// Execution environment of an enqueue: the stream, the buffers and the
// configuration the launch runs with.
struct Exec {
    hipStream_t s = nullptr;
    Buffers* b = nullptr;
    cpv_config cfg{};
    cpv_mlp* predictor = nullptr;
    cpv_context* ctx = nullptr;   // CU count and the occupancy cache
    bool frozen = false;          // recording a graph: no allocation
    hipEvent_t r0 = nullptr, r1 = nullptr;
    uint32_t* overflow = nullptr; // overflow word shared by several launches (nullptr: the launch's own)
    unsigned long long* fragments = nullptr;   // covered-fragment counter (nullptr: the launch's own)
    const uint32_t* fixed_order = nullptr;     // learned tile order computed earlier (nullptr: computed here)
};

// Exclusive scan of n counts into out (out[n] = the total when write_total);
// total64, when given, receives the total in 64 bits.
int scan_u32(Exec& x, const uint32_t* in, uint32_t n, uint32_t* out, bool write_total, uint64_t* total64 = nullptr)
{
    if (n == 0) {
        if (write_total) fill_u32(x.s, out, 0u, 1);
        if (total64) fill_u32(x.s, total64, 0u, 2);
        return CPV_OK;
    }
    const uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    if (x.b->sums.ensure((size_t)blocks * 8, x.frozen)) return CPV_ERROR_OUT_OF_MEMORY;
    uint64_t* sums = x.b->sums.as<uint64_t>();
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(blocks), dim3(kScanThreads), 0, x.s, in, n, sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, x.s, sums, blocks, total64);
    hipLaunchKernelGGL(scan_down_kernel, dim3(blocks), dim3(kScanThreads), 0, x.s, in, n, (const uint64_t*)sums, out,
                       write_total ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return CPV_OK;
}

// Stable LSD radix sort of (key, value) pairs by the low `bits` key bits;
// n is read on the device.  Returns the buffer index holding the result.
int radix_sort(Exec& x, DevBuf* keys, DevBuf* vals, const uint32_t* d_n, int bits, int* result)
{
    Buffers& b = *x.b;
    int src = 0;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(radix_hist_kernel, dim3(kSortSegments), dim3(kSortThreads), 0, x.s,
                           (const uint32_t*)keys[src].p, d_n, shift, b.hist.as<uint32_t>());
        int rc = scan_u32(x, b.hist.as<uint32_t>(), 256u * kSortSegments, b.hist_base.as<uint32_t>(), false);
        if (rc) return rc;
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(kSortSegments), dim3(kSortThreads), 0, x.s,
                           (const uint32_t*)keys[src].p, (const uint32_t*)vals[src].p, keys[1 - src].as<uint32_t>(),
                           vals[1 - src].as<uint32_t>(), d_n, shift, (const uint32_t*)b.hist_base.p);
        HIP_TRY(hipGetLastError());
        src = 1 - src;
    }
    *result = src;
    return CPV_OK;
}

int effective_scheduler(const Exec& x)
{
    if (x.cfg.scheduler == CPV_SCHED_PREDICTED && !x.predictor) return CPV_SCHED_COUNTER;
    return (int)x.cfg.scheduler;
}

bool needs_features(const Exec& x) { return effective_scheduler(x) == CPV_SCHED_PREDICTED || x.cfg.record_costs; }

int resident_workgroups(Exec& x, const KernelRef& k, int block)
{
    const void* key = k.host_fn ? k.host_fn : (const void*)k.module_fn;
    auto it = x.ctx->occupancy.find(key);
    int per_cu = 0;
    if (it != x.ctx->occupancy.end()) {
        per_cu = it->second;
    } else {
        if (k.host_fn) {
            HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k.host_fn, block, 0));
        } else {
            HIP_TRY(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k.module_fn, block, 0));
        }
        per_cu = std::max(per_cu, 1);
        x.ctx->occupancy[key] = per_cu;
    }
    return per_cu * std::max(x.ctx->cu_count, 1);
}

// Slab sums of the scans of a launch: the primitive counts' and the radix
// sorts' histogram scans (256 digits x kSortSegments segments).
size_t scan_sums_bytes(uint32_t prims)
{
    const size_t most = std::max<size_t>(prims, (size_t)256 * kSortSegments);
    return (most + kScanTile - 1) / kScanTile * 8;
}

// Buffers whose size depends only on the plan.
int ensure_front(Exec& x, const LaunchPlan& P)
{
    Buffers& b = *x.b;
    const bool f = x.frozen;
    if (b.vout.ensure((size_t)std::max(P.vout_total, 1u) * P.stride * 4, f) ||
        b.recs.ensure((size_t)std::max(P.prim_total, 1u) * sizeof(TriSetup), f) ||
        b.counts.ensure((size_t)std::max(P.prim_total, 1u) * 4, f) || b.offsets.ensure(((size_t)P.prim_total + 1) * 4, f) ||
        b.stats.ensure(kStatAlloc * 4, f) || b.sums.ensure(scan_sums_bytes(P.prim_total), f))
        return CPV_ERROR_OUT_OF_MEMORY;
    if (P.lines && b.line_ext.ensure((size_t)std::max(P.prim_total, 1u) * sizeof(LineExt), f)) return CPV_ERROR_OUT_OF_MEMORY;
    return CPV_OK;
}

// Buffers sized by the pair capacity and the tile count.
int ensure_back(Exec& x, const LaunchPlan& P, uint32_t pair_cap)
{
    Buffers& b = *x.b;
    const bool f = x.frozen;
    const size_t pairs = (size_t)std::max(pair_cap, 1u) * 4;
    const size_t tiles = (size_t)std::max(P.tiles, 1u);
    if (b.keys[0].ensure(pairs, f) || b.keys[1].ensure(pairs, f) || b.vals[0].ensure(pairs, f) || b.vals[1].ensure(pairs, f) ||
        b.tile_start.ensure((tiles + 1) * 4, f) || b.hist.ensure((size_t)256 * kSortSegments * 4, f) ||
        b.hist_base.ensure(((size_t)256 * kSortSegments + 1) * 4, f) || b.big.ensure(((size_t)P.prim_total + 1) * 4, f))
        return CPV_ERROR_OUT_OF_MEMORY;
    if (needs_features(x)) {
        if (b.features.ensure(tiles * kFeatureStride * 4, f) || b.pred.ensure(tiles * 4, f) || b.okeys[0].ensure(tiles * 4, f) ||
            b.okeys[1].ensure(tiles * 4, f) || b.ovals[0].ensure(tiles * 4, f) || b.ovals[1].ensure(tiles * 4, f))
            return CPV_ERROR_OUT_OF_MEMORY;
    }
    if (x.cfg.record_costs && b.traces.ensure(tiles * x.cfg.waves_per_tile * 4, f)) return CPV_ERROR_OUT_OF_MEMORY;
    return CPV_OK;
}

// Vertex stage, triangle setup and the pair-count scan; the pair total is
// left at offsets[prim_total].
int enqueue_front(Exec& x, const LaunchPlan& P, uint8_t* arena)
{
    int rc = ensure_front(x, P);
    if (rc) return rc;
    Buffers& b = *x.b;
    const DrawDesc* d_dd = (const DrawDesc*)(arena + P.o_dd);
    const TargetParams* d_tp = (const TargetParams*)(arena + P.o_tp);
    if (P.vout_total > 0) {
        VertexArgs va{d_dd, (const uint32_t*)(arena + P.o_vb), b.vout.as<float>(), P.stride, 0};
        HIP_TRY(launch(P.prog.vertex, dim3((uint32_t)P.vblock.size()), dim3(256), &va, x.s));
    }
    if (P.prim_total > 0) {
        SetupArgs sa{d_dd, (const uint32_t*)(arena + P.o_sb), d_tp, b.vout.as<float>(), b.recs.as<TriSetup>(),
                     P.lines ? b.line_ext.as<LineExt>() : nullptr, b.counts.as<uint32_t>(), P.stride, P.tile_shift};
        hipLaunchKernelGGL(setup_kernel, dim3((uint32_t)P.sblock.size()), dim3(256), 0, x.s, sa);
        HIP_TRY(hipGetLastError());
    }
    return scan_u32(x, b.counts.as<uint32_t>(), P.prim_total, b.offsets.as<uint32_t>(), true,
                    (uint64_t*)(b.stats.as<uint32_t>() + kStatPairTotal));
}

// Where the binning stage left its results for the raster kernel.
struct Binned {
    int src = 0;                       // pair buffer holding the tile-sorted pairs
    const uint32_t* order = nullptr;   // learned tile order (nullptr: index order)
};

// Emit, sort, tile ranges and scheduling, with the pair arrays holding at
// most pair_cap pairs.  Touches no attachment: a recorded batch runs it for
// one submit while the previous submit's raster kernel is still running.
int enqueue_bin(Exec& x, const LaunchPlan& P, uint8_t* arena, uint32_t pair_cap, Binned* out)
{
    int rc = ensure_back(x, P, pair_cap);
    if (rc) return rc;
    Buffers& b = *x.b;
    uint32_t* stats = b.stats.as<uint32_t>();
    const DrawDesc* d_dd = (const DrawDesc*)(arena + P.o_dd);
    const FragmentState* d_fs = (const FragmentState*)(arena + P.o_fs);
    const TargetParams* d_tp = (const TargetParams*)(arena + P.o_tp);
    const uint32_t cap = std::max(pair_cap, 1u);
    hipLaunchKernelGGL(clamp_pairs_kernel, dim3(1), dim3(64), 0, x.s, (const uint64_t*)(stats + kStatPairTotal), cap,
                       stats + kStatPairs, x.overflow ? x.overflow : stats + kStatOverflow);
    // emit pairs in primitive order
    if (P.prim_total) {
        fill_u32(x.s, b.big.p, 0u, 1);
        EmitArgs ea{b.recs.as<TriSetup>(), b.counts.as<uint32_t>(), b.offsets.as<uint32_t>(), d_dd, d_tp, b.keys[0].as<uint32_t>(),
                    b.vals[0].as<uint32_t>(), b.big.as<uint32_t>(), P.prim_total, P.tile_shift, cap};
        const uint32_t blocks = std::min<uint32_t>((P.prim_total + 255) / 256, 8192);
        hipLaunchKernelGGL(emit_kernel, dim3(blocks), dim3(256), 0, x.s, ea);
        hipLaunchKernelGGL(emit_big_kernel, dim3(1024), dim3(256), 0, x.s, ea);
        HIP_TRY(hipGetLastError());
    }
    // stable radix sort of the pairs by tile
    int src = 0;
    if (P.tiles > 1) {
        int bits = 0;
        while ((1u << bits) < P.tiles) bits++;
        rc = radix_sort(x, b.keys, b.vals, stats + kStatPairs, bits, &src);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(tile_ranges_kernel, dim3(P.tiles / 256 + 1), dim3(256), 0, x.s, (const uint32_t*)b.keys[src].p,
                       (const uint32_t*)(stats + kStatPairs), P.tiles, b.tile_start.as<uint32_t>());
    HIP_TRY(hipGetLastError());

    // scheduling (a recorded batch computes its learned order once, when it
    // is recorded, and its submits reuse it)
    const int sched = effective_scheduler(x);
    const uint32_t* order = nullptr;
    const bool reuse_order = sched == CPV_SCHED_PREDICTED && x.fixed_order;
    if (needs_features(x) && P.tiles && !(reuse_order && !x.cfg.record_costs)) {
        FeatureArgs fa{b.recs.as<TriSetup>(), b.vals[src].as<uint32_t>(), b.tile_start.as<uint32_t>(), d_tp, d_dd, d_fs,
                       b.features.as<float>(), P.tiles, (uint32_t)P.targets.size(), (int32_t)x.cfg.tile_size,
                       (int32_t)x.cfg.waves_per_tile, P.hiz_op >= 0 ? 1 : 0, 0};
        hipLaunchKernelGGL(tile_features_kernel, dim3((P.tiles + 3) / 4), dim3(256), 0, x.s, fa);
        HIP_TRY(hipGetLastError());
    }
    if (reuse_order && P.tiles) {
        order = x.fixed_order;
    } else if (sched == CPV_SCHED_PREDICTED && P.tiles) {
        rc = mlp_enqueue_predict(x.predictor, b.features.as<float>(), P.tiles, b.pred.as<float>(), x.s);
        if (rc) return rc;
        fill_u32(x.s, stats + kStatPredMax, 0u, 1);
        hipLaunchKernelGGL(pred_max_kernel, dim3(std::min<uint32_t>((P.tiles + 255) / 256, 256)), dim3(256), 0, x.s,
                           (const float*)b.pred.p, P.tiles, stats + kStatPredMax);
        hipLaunchKernelGGL(heavy_threshold_kernel, dim3(1), dim3(64), 0, x.s, (const uint32_t*)(stats + kStatPredMax),
                           (float*)(stats + kStatPrio));
        hipLaunchKernelGGL(order_keys_kernel, dim3((P.tiles + 255) / 256), dim3(256), 0, x.s, (const float*)b.pred.p,
                           (const float*)(stats + kStatPrio), P.tiles, b.okeys[0].as<uint32_t>(), b.ovals[0].as<uint32_t>());
        HIP_TRY(hipGetLastError());
        fill_u32(x.s, stats + kStatTiles, P.tiles, 1);
        int osrc = 0;
        rc = radix_sort(x, b.okeys, b.ovals, stats + kStatTiles, 16, &osrc);
        if (rc) return rc;
        order = b.ovals[osrc].as<uint32_t>();
    }
    out->src = src;
    out->order = order;
    return CPV_OK;
}

// The raster kernel of a launch whose binning enqueue_bin enqueued.
int enqueue_raster(Exec& x, const LaunchPlan& P, uint8_t* arena, const Binned& bn)
{
    Buffers& b = *x.b;
    uint32_t* stats = b.stats.as<uint32_t>();
    const DrawDesc* d_dd = (const DrawDesc*)(arena + P.o_dd);
    const FragmentState* d_fs = (const FragmentState*)(arena + P.o_fs);
    const TargetParams* d_tp = (const TargetParams*)(arena + P.o_tp);
    const int sched = effective_scheduler(x);
    const int src = bn.src;
    const uint32_t* order = bn.order;
    const KernelRef& rk = P.prog.raster[P.raster_variant];
    const int block = 64 * (int)x.cfg.waves_per_tile;
    RasterArgs ra{};
    ra.recs = b.recs.as<TriSetup>();
    ra.tile_prims = b.vals[src].as<uint32_t>();
    ra.tile_start = b.tile_start.as<uint32_t>();
    ra.targets = d_tp;
    ra.draws = d_dd;
    ra.fstates = d_fs;
    ra.vout = b.vout.as<float>();
    ra.line_ext = P.lines ? b.line_ext.as<LineExt>() : nullptr;
    ra.order = order;
    ra.queue = stats + kStatQueues;
    ra.fragments = x.fragments ? x.fragments : (unsigned long long*)(stats + kStatFragments);
    ra.overflow = x.overflow ? x.overflow : stats + kStatOverflow;
    ra.trace_cycles = x.cfg.record_costs ? b.traces.as<float>() : nullptr;
    ra.vout_stride = P.stride;
    ra.tiles = P.tiles;
    ra.target_count = (uint32_t)P.targets.size();
    // The counter-ordered queue is a persistent kernel; the learned order
    // is handed to the hardware dispatcher (workgroup i takes the i-th tile
    // of the order), which starts the predicted-heavy tiles first with no
    // queue atomics.  (Raising the heavy tiles' wave priority with
    // s_setprio was measured on the 4K skewed frame: it made the raster
    // kernel 2.8 -> 3.3 ms, 22% more wave cycles in total, with the heavy
    // tiles themselves no faster.)
    ra.persistent = sched == CPV_SCHED_COUNTER ? 1u : 0u;
    ra.hiz_op = P.hiz_op;
    ra.hiz_depth_kind = P.hiz_kind;
    ra.uniform_fstate = P.uniform_fstate;
    uint32_t grid = P.tiles;
    if (ra.persistent) {
        const int resident = resident_workgroups(x, rk, block);
        if (resident <= 0) return CPV_ERROR_DEVICE;
        grid = std::min<uint32_t>(P.tiles, (uint32_t)resident);
        fill_u32(x.s, stats + kStatQueues, 0u, kRasterQueues * kRasterQueueStride);
    }
    if (grid == 0) return CPV_OK;
    if (x.r0) HIP_TRY(hipEventRecord(x.r0, x.s));
    HIP_TRY(launch(rk, dim3(grid), dim3(block), &ra, x.s));
    if (x.r1) HIP_TRY(hipEventRecord(x.r1, x.s));
    return CPV_OK;
}

int enqueue_back(Exec& x, const LaunchPlan& P, uint8_t* arena, uint32_t pair_cap)
{
    Binned bn;
    const int rc = enqueue_bin(x, P, arena, pair_cap, &bn);
    return rc ? rc : enqueue_raster(x, P, arena, bn);
}

// Uploads staged descriptors through the context's pinned staging ring.
int upload_staging(cpv_context* c, DevBuf& arena, const Staging& st) { return stage_upload(c, arena.p, st.bytes.data(), st.bytes.size()); }

Exec context_exec(cpv_context* c, Buffers* b)
{
    Exec x;
    x.s = c->stream;
    x.b = b;
    x.cfg = c->cfg;
    x.predictor = c->predictor;
    x.ctx = c;
    return x;
}

}  // namespace

// An immediate launch with what it takes to run it again.
struct cpvrt::PendingLaunch {
    LaunchPlan P;
    Staging st;
    cpv_config cfg{};
    uint32_t set = 0;            // the context's buffer set it runs in
    bool in_scratch = false;     // ran in the scratch array (cpv_render_pass_noalloc)
    uint32_t scratch_cap = 0;    // its pair capacity there
};

void cpvrt::destroy_pending(PendingLaunch* p) { delete p; }

namespace {

// Enqueues an immediate launch in its buffer set with the context's pair
// capacity (at least 16 pairs per primitive and 1M): the descriptors'
// upload, the front half (vertex, setup, binning) and the read-back of the
// 64-bit pair total and overflow word on front_stream, then the raster
// kernel on the context's stream after it.  The next call on the context
// checks the pair count (resolve_pending): no host round trip sits between
// a launch's binning and its raster kernel.  The front half waits for the
// raster kernel that last used the set and, unless the previous context
// call was a launch too, for all work enqueued on the context's stream; so
// back-to-back launches run one front half under the previous raster.
int enqueue_checked(cpv_context* c, PendingLaunch& L)
{
    Buffers& B = c->lbuf[L.set];
    hipStream_t fs = c->front_stream;
    const bool chained = c->chain_seq == c->enqueue_seq;
    if (!chained) {
        HIP_TRY(hipEventRecord(c->ev_pre, c->stream));
        HIP_TRY(hipStreamWaitEvent(fs, c->ev_pre, 0));
    }
    if (c->set_used[L.set]) HIP_TRY(hipStreamWaitEvent(fs, c->set_raster_done[L.set], 0));
    if (c->call_first_launch) {
        HIP_TRY(hipEventRecord(c->ev_begin, fs));
        c->call_first_launch = false;
    }
    if (B.arena.ensure(L.st.bytes.size())) return CPV_ERROR_OUT_OF_MEMORY;
    uint8_t* arena = B.arena.as<uint8_t>();
    patch_plan(L.P, L.st, arena);
    int rc = stage_upload(c, arena, L.st.bytes.data(), L.st.bytes.size(), true, fs);
    if (rc) return rc;
    const bool fresh = B.stats.p == nullptr;
    if (B.stats.ensure(kStatAlloc * 4)) return CPV_ERROR_OUT_OF_MEMORY;
    if (fresh) fill_u32(fs, B.stats.p, 0u, kStatAlloc);   // the overflow word starts clear
    Exec x = context_exec(c, &B);
    x.cfg = L.cfg;
    x.s = fs;
    x.fragments = (unsigned long long*)(c->buf.stats.as<uint32_t>() + kStatFragments);
    rc = enqueue_front(x, L.P, arena);
    if (rc) return rc;
    const uint64_t want = std::max<uint64_t>(c->pair_cap, 16ull * L.P.prim_total + (1u << 20));
    const uint32_t cap = (uint32_t)std::min<uint64_t>(want, UINT32_MAX);
    Binned bn;
    rc = enqueue_bin(x, L.P, arena, cap, &bn);
    if (rc) return rc;
    c->pair_cap = std::max<uint64_t>(c->pair_cap, cap);
    uint32_t* stats = B.stats.as<uint32_t>();
    HIP_TRY(hipMemcpyAsync(c->readback + 32, stats + kStatPairTotal, 8, hipMemcpyDeviceToHost, fs));
    HIP_TRY(hipMemcpyAsync(c->readback + 34, stats + kStatOverflow, 4, hipMemcpyDeviceToHost, fs));
    HIP_TRY(hipEventRecord(c->ev_check, fs));
    HIP_TRY(hipEventRecord(c->set_front_done[L.set], fs));
    HIP_TRY(hipStreamWaitEvent(c->stream, c->set_front_done[L.set], 0));
    x.s = c->stream;
    x.r0 = c->ev_raster0;
    x.r1 = c->ev_raster1;
    rc = enqueue_raster(x, L.P, arena, bn);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(c->set_raster_done[L.set], c->stream));
    c->set_used[L.set] = true;
    c->chain_seq = c->enqueue_seq;
    return CPV_OK;
}

}  // namespace

// ------------------------------------------------ allocation-free launches
//
// cpv_render_pass_noalloc runs a render pass with every working buffer in
// one file-scope device array per GPU: the control words, the descriptors,
// the vertex outputs, the triangle records, the scans, the tile ranges and
// the four pair arrays.  On the host the array's name is no device address;
// host code uses only the address hipGetSymbolAddress returns for it with
// the device current (cpv_context::scratch, taken when the context is
// created).  Every call on a GPU shares that GPU's array: calls that overlap
// on one GPU, from two contexts or two threads, race.

constexpr size_t kScratchBytes = 64ull << 20;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

int cpvrt::scratch_setup(cpv_context* c)
{
    static std::mutex lock;
    static std::map<int, uint8_t*> by_device;
    std::lock_guard<std::mutex> g(lock);
    auto it = by_device.find(c->device);
    if (it == by_device.end()) {
        HIP_TRY(hipSetDevice(c->device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(CPV_ERROR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        it = by_device.emplace(c->device, (uint8_t*)p).first;
    }
    c->scratch = it->second;
    return CPV_OK;
}

size_t cpvrt::scratch_bytes() { return kScratchBytes; }

namespace {

// Places every buffer of launch P in the scratch array at base, with the
// sizes ensure_front and ensure_back ask for: the control words first (at
// the same address for every launch), then the descriptors, vertex outputs,
// records, scans and tile ranges; the four pair arrays share what is left
// and *pair_cap receives their capacity.  Touches no memory.
int carve_scratch(uint8_t* base, const LaunchPlan& P, size_t arena_bytes, Buffers* b, uint32_t* pair_cap)
{
    size_t at = 0;
    auto take = [&](DevBuf& d, size_t bytes) {
        at = (at + 255) & ~(size_t)255;
        d.p = base + at;
        d.cap = bytes;
        at += bytes;
    };
    const size_t prims = std::max(P.prim_total, 1u), tiles = std::max(P.tiles, 1u);
    take(b->stats, kStatAlloc * 4);
    take(b->arena, std::max<size_t>(arena_bytes, 256));
    take(b->vout, (size_t)std::max(P.vout_total, 1u) * P.stride * 4);
    take(b->recs, prims * sizeof(TriSetup));
    take(b->counts, prims * 4);
    take(b->offsets, ((size_t)P.prim_total + 1) * 4);
    take(b->sums, scan_sums_bytes(P.prim_total));
    if (P.lines) take(b->line_ext, prims * sizeof(LineExt));
    take(b->tile_start, (tiles + 1) * 4);
    take(b->hist, (size_t)256 * kSortSegments * 4);
    take(b->hist_base, ((size_t)256 * kSortSegments + 1) * 4);
    take(b->big, ((size_t)P.prim_total + 1) * 4);
    // each pair array starts on a 256-byte boundary: at most 4 x 255 bytes of padding
    const size_t left = at + 1024 < kScratchBytes ? kScratchBytes - at - 1024 : 0;
    const uint64_t cap = std::min<uint64_t>(left / 16, UINT32_MAX);
    if (cap == 0)
        return fail(CPV_ERROR_UNSUPPORTED,
                    "a launch of %u primitives over %u tiles needs more than the %zu MiB of the allocation-free entry point",
                    P.prim_total, P.tiles, kScratchBytes >> 20);
    for (DevBuf* d : {&b->keys[0], &b->keys[1], &b->vals[0], &b->vals[1]}) take(*d, (size_t)cap * 4);
    *pair_cap = (uint32_t)cap;
    return CPV_OK;
}

// Enqueues launch L with every buffer in the scratch array: its descriptors
// through a staging slot that holds them without growing, then the front
// half, the binning with the pair capacity left in the array, and the raster
// kernel, then the read-back of its pair total and overflow word, which the
// next call on the context checks.  Allocates nothing.
int enqueue_scratch(cpv_context* c, PendingLaunch& L)
{
    Buffers b;
    uint32_t cap = 0;
    int rc = carve_scratch(c->scratch, L.P, L.st.bytes.size(), &b, &cap);
    if (rc) return rc;
    uint8_t* arena = b.arena.as<uint8_t>();
    patch_plan(L.P, L.st, arena);
    rc = stage_upload(c, arena, L.st.bytes.data(), L.st.bytes.size(), false);
    if (rc) return rc;
    Exec x = context_exec(c, &b);
    x.cfg = L.cfg;
    x.cfg.record_costs = 0;
    x.predictor = nullptr;   // the learned order needs the predictor's buffers: the counter queue runs instead
    x.frozen = true;         // a buffer that does not fit is an error, never an allocation
    rc = enqueue_front(x, L.P, arena);
    if (rc) return rc;
    x.r0 = c->ev_raster0;
    x.r1 = c->ev_raster1;
    rc = enqueue_back(x, L.P, arena, cap);
    if (rc) return rc;
    L.in_scratch = true;
    L.scratch_cap = cap;
    uint32_t* stats = b.stats.as<uint32_t>();
    HIP_TRY(hipMemcpyAsync(c->readback + 32, stats + kStatPairTotal, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(c->readback + 34, stats + kStatOverflow, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_check, c->stream));
    return CPV_OK;
}

}  // namespace

// Checks the pending launch's pair count; a launch whose pairs did not fit
// drew nothing (its raster kernel saw the overflow word) and is run again
// here with arrays of the exact size, before any later work is enqueued.
// Without may_allocate the launch runs again in the scratch array, if its
// pairs fit there; a launch that already ran there cannot grow.
int cpvrt::resolve_pending(cpv_context* c, bool may_allocate)
{
    if (c->pending_batch) {
        cpv_batch* B = c->pending_batch;
        const uint32_t si = c->pending_set;
        c->pending_batch = nullptr;
        HIP_TRY(hipSetDevice(c->device));
        HIP_TRY(hipEventSynchronize(c->ev_check));
        if (!c->readback[36]) {
            c->fences_after_pending.clear();
            return CPV_OK;
        }
        const int rc = batch_redo(c, B, si, may_allocate);
        if (rc) {
            c->fences_after_pending.clear();
            return rc;
        }
        rerecord_fences(c);
        return resolve_pending(c, may_allocate);
    }
    if (!c->pending) {
        c->fences_after_pending.clear();
        return CPV_OK;
    }
    std::unique_ptr<PendingLaunch> L(c->pending);
    c->pending = nullptr;
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventSynchronize(c->ev_check));
    uint64_t total = 0;
    std::memcpy(&total, c->readback + 32, 8);
    if (!c->readback[34]) {
        c->fences_after_pending.clear();
        return CPV_OK;
    }
    int rc = CPV_OK;
    if (total > (uint64_t)UINT32_MAX)
        rc = fail(CPV_ERROR_OVERFLOW,
                  "a render call's primitives need %llu (tile, primitive) pairs, more than 2^32 - 1: split its draws "
                  "or use larger tiles (nothing was drawn)",
                  (unsigned long long)total);
    else if (L->in_scratch)
        rc = fail(CPV_ERROR_OVERFLOW,
                  "a render call's primitives need %llu (tile, primitive) pairs; the allocation-free entry point had room "
                  "for %u (nothing was drawn from that launch on)",
                  (unsigned long long)total, L->scratch_cap);
    else if (may_allocate) {
        c->pair_cap = std::min<uint64_t>(total + total / 8 + 4096, UINT32_MAX);
        fill_u32(c->stream, c->lbuf[L->set].stats.as<uint32_t>() + kStatOverflow, 0u, 1);
        c->chain_seq = ~0ull;   // the redraw's front half waits for the context stream (the fill, the skipped raster)
        rc = enqueue_checked(c, *L);
    } else {
        Buffers probe;
        uint32_t cap = 0;
        rc = carve_scratch(c->scratch, L->P, L->st.bytes.size(), &probe, &cap);
        if (rc == CPV_OK && cap < total)
            rc = fail(CPV_ERROR_OVERFLOW,
                      "the previous render call needs %llu (tile, primitive) pairs, more than the allocation-free entry "
                      "point holds (%u): call cpv_synchronize after it (nothing was drawn)",
                      (unsigned long long)total, cap);
        if (rc == CPV_OK) {
            fill_u32(c->stream, c->scratch, 0u, kStatWords);
            rc = enqueue_scratch(c, *L);
        }
        c->last_traces = false;
        c->last_stats = (uint32_t*)c->scratch;
    }
    if (rc) {
        c->fences_after_pending.clear();
        return rc;
    }
    c->launches++;
    rerecord_fences(c);
    c->pending = L.release();
    return resolve_pending(c, may_allocate);   // the exact capacity holds every pair: the check passes
}

namespace {

// Byte ranges of device memory, for the hazards between render passes.

void attachment_spans(const cpv_pass_desc& p, std::vector<Span>* out)
{
    auto add = [&](const cpv_attachment_desc& a) {
        FormatDesc f;
        if (!a.address || !format_desc(a.format, &f)) return;
        out->push_back({a.address, a.address + (uint64_t)p.width * p.height * (uint64_t)f.total});
    };
    for (uint32_t k = 0; k < p.color_count && k < CPV_MAX_COLOR_ATTACHMENTS; k++) add(p.color[k]);
    if (p.has_depth_stencil) add(p.depth_stencil);
}

void texture_spans(const cpv_draw_desc& d, std::vector<Span>* out)
{
    for (const cpv_texture_desc& t : d.textures) {
        FormatDesc f;
        if (!t.address || !format_desc(t.format, &f)) continue;
        // every level of a mip chain fits in twice the base level
        out->push_back({t.address, t.address + 2ull * t.width * t.height * (uint64_t)f.total});
    }
}

bool spans_meet(const std::vector<Span>& a, const std::vector<Span>& b)
{
    for (const Span& x : a)
        for (const Span& y : b)
            if (x.meets(y)) return true;
    return false;
}

// One immediate launch on the context: plan, then upload, front half and
// back half with no host synchronisation (enqueue_checked).
int run_launch(cpv_context* c, const cpv_config& cfg, const Program& prog, const std::vector<TargetParams>& targets,
               const std::vector<LaunchDraw>& lds, const std::vector<cpv_pass_desc>& passes,
               const std::vector<uint32_t>& draw_target)
{
    int rc = resolve_pending(c);   // the previous launch's pairs, before this one reuses the buffers
    if (rc) return rc;
    std::unique_ptr<PendingLaunch> L(new PendingLaunch());
    L->cfg = cfg;
    rc = plan_launch(cfg, prog, targets, lds, passes, draw_target, &L->P);
    if (rc) return rc;
    stage_plan(L->P, L->st);
    L->set = c->next_set;
    c->next_set ^= 1u;
    // The front half may run before the previous launch's raster kernel has
    // finished: not when this launch's vertex stage reads memory that
    // launch draws into (an attachment used as a vertex, index or uniform
    // buffer).
    std::vector<Span> reads;
    for (const LaunchDraw& d : lds) {
        const cpv_draw_desc& dd = *d.draw;
        for (int b = 0; b < CPV_MAX_VERTEX_BINDINGS; b++)
            if (dd.vertex_address[b]) reads.push_back({dd.vertex_address[b], dd.vertex_address[b] + dd.vertex_bytes[b]});
        if (dd.indexed && dd.index_address)
            reads.push_back({dd.index_address, dd.index_address + ((uint64_t)dd.first + dd.count) * dd.index_size});
        for (int u = 0; u < CPV_MAX_UNIFORMS; u++)
            if (dd.uniform_address[u]) reads.push_back({dd.uniform_address[u], dd.uniform_address[u] + (64u << 10)});
    }
    if (spans_meet(reads, c->last_written)) c->chain_seq = ~0ull;
    c->last_written.clear();
    for (const cpv_pass_desc& p : passes) attachment_spans(p, &c->last_written);
    rc = enqueue_checked(c, *L);
    if (rc) return rc;
    c->last_set = L->set;
    const LaunchPlan& P = L->P;
    c->last_tiles = P.tiles;
    c->last_waves = cfg.waves_per_tile;
    c->last_tile_size = cfg.tile_size;
    c->last_traces = cfg.record_costs != 0;
    c->last_target_base.clear();
    for (const TargetParams& t : P.targets) c->last_target_base.push_back(t.tile_base);
    c->pending = L.release();
    c->launches++;
    return CPV_OK;
}

// Points and lines raster in the 16x16 tile, four-wave kernel variant that
// carries them: a call with any such draw runs every launch in that
// configuration.
cpv_config call_config(const cpv_config& cfg, bool points_or_lines)
{
    cpv_config e = cfg;
    if (points_or_lines) {
        e.tile_size = 16;
        e.waves_per_tile = 4;
    }
    return e;
}

bool jobs_have_points_or_lines(const cpv_job_desc* jobs, uint32_t n)
{
    for (uint32_t j = 0; j < n; j++)
        if (jobs[j].pipeline.topology < CPV_TOPOLOGY_TRIANGLE_LIST) return true;
    return false;
}

bool same_program(const cpv_pipeline_desc& a, const cpv_pipeline_desc& b)
{
    if (a.program || b.program) return a.program == b.program;
    return a.vs_kind == b.vs_kind && a.fs_kind == b.fs_kind;
}

// One launch of a frame: a shader program and the draws of one or more
// independent render targets, whose tile ranges are laid end to end.
struct LaunchSpec {
    Program prog;
    std::vector<TargetParams> targets;
    std::vector<cpv_pass_desc> passes;
    std::vector<LaunchDraw> lds;
    std::vector<uint32_t> dt;
};

// Groups jobs by shader program: one launch per group, every job its own
// render target.
int group_jobs(const cpv_job_desc* jobs, uint32_t job_count, int tile, std::vector<LaunchSpec>* groups)
{
    std::vector<bool> done(job_count, false);
    for (uint32_t g = 0; g < job_count; g++) {
        if (done[g]) continue;
        LaunchSpec G;
        int rc = resolve_program(jobs[g].pipeline, &G.prog);
        if (rc) return rc;
        uint32_t tile_base = 0;
        for (uint32_t j = g; j < job_count; j++) {
            if (done[j] || !same_program(jobs[j].pipeline, jobs[g].pipeline)) continue;
            done[j] = true;
            TargetParams t;
            rc = build_target(jobs[j].pass, true, tile, tile_base, &t);
            if (rc) return rc;
            if ((uint64_t)tile_base + (uint64_t)t.tiles_x * t.tiles_y >= (1ull << 31))
                return fail(CPV_ERROR_UNSUPPORTED, "more than 2^31 tiles in one launch");
            tile_base += t.tiles_x * t.tiles_y;
            const uint32_t ti = (uint32_t)G.targets.size();
            G.targets.push_back(t);
            G.passes.push_back(jobs[j].pass);
            for (uint32_t inst = 0; inst < jobs[j].draw.instance_count; inst++) {
                G.lds.push_back({&jobs[j].pipeline, &jobs[j].draw, inst});
                G.dt.push_back(ti);
            }
        }
        groups->push_back(std::move(G));
    }
    return CPV_OK;
}

}  // namespace

// ------------------------------------------------------------ render calls

namespace {

// Opens an immediate render call: checks the previous launch, keeps the
// launch chain when the previous context call was a render call too, clears
// the call's fragment counter (on the context stream, before this call's
// raster kernels) and arms the begin event for its first launch.
int begin_render_call(cpv_context* c)
{
    const bool chained = c->chain_seq == c->enqueue_seq;
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    if (chained) c->chain_seq = c->enqueue_seq;
    HIP_TRY(hipSetDevice(c->device));
    if (c->buf.stats.ensure(kStatAlloc * 4)) return CPV_ERROR_OUT_OF_MEMORY;
    fill_u32(c->stream, c->buf.stats.p, 0u, kStatWords);
    c->call_first_launch = true;
    return CPV_OK;
}

int end_render_call(cpv_context* c)
{
    if (c->call_first_launch) {   // nothing was launched
        HIP_TRY(hipEventRecord(c->ev_begin, c->stream));
        c->call_first_launch = false;
    }
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    c->timing_valid = true;
    c->last_count_exact = c->cfg.count_fragments != 0;
    c->last_stats = c->buf.stats.as<uint32_t>();
    return CPV_OK;
}

}  // namespace

namespace {

// The launches of one render pass: one per run of consecutive draws that
// share a shader program (the first applies the pass's load ops, the rest
// load what it stored); a pass with no draw is one empty launch, which still
// applies the load ops to every tile.
struct PassLaunch {
    Program prog;
    std::vector<LaunchDraw> lds;
    std::vector<uint32_t> dt;
    bool first = true;
};

int split_pass(const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws, const uint32_t* draw_pipeline, uint32_t draw_count,
               std::vector<PassLaunch>* out)
{
    if (draw_count == 0) {
        PassLaunch L;
        if (!builtin_program(CPV_VS_PASSTHROUGH, CPV_FS_COLOR, &L.prog)) return fail(CPV_ERROR_DEVICE, "built-in table");
        out->push_back(std::move(L));
        return CPV_OK;
    }
    uint32_t i = 0;
    while (i < draw_count) {
        PassLaunch L;
        L.first = out->empty();
        const uint32_t pi0 = draw_pipeline ? draw_pipeline[i] : 0;
        int rc = resolve_program(pipelines[pi0], &L.prog);
        if (rc) return rc;
        uint32_t j = i;
        while (j < draw_count) {
            const uint32_t pj = draw_pipeline ? draw_pipeline[j] : 0;
            if (!same_program(pipelines[pj], pipelines[pi0])) break;
            for (uint32_t inst = 0; inst < draws[j].instance_count; inst++) {
                L.lds.push_back({&pipelines[pj], &draws[j], inst});
                L.dt.push_back(0);
            }
            j++;
        }
        i = j;
        out->push_back(std::move(L));
    }
    return CPV_OK;
}

int check_pass_args(cpv_context* c, const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws,
                    uint32_t draw_count)
{
    if (!c || !pass || (draw_count && (!pipelines || !draws))) return fail(CPV_ERROR_INVALID, "null argument");
    if (pass->color_count > CPV_MAX_COLOR_ATTACHMENTS) return fail(CPV_ERROR_UNSUPPORTED, "%u colour attachments", pass->color_count);
    return CPV_OK;
}

bool pass_has_points_or_lines(const cpv_pipeline_desc* pipelines, const uint32_t* draw_pipeline, uint32_t draw_count)
{
    bool any = false;
    for (uint32_t k = 0; k < draw_count; k++)
        any = any || pipelines[draw_pipeline ? draw_pipeline[k] : 0].topology < CPV_TOPOLOGY_TRIANGLE_LIST;
    return any;
}

// The launches of several render passes, each with its own draws
// (draws first_draw[p] .. first_draw[p + 1] - 1 belong to pass p).  The
// result is that of running the passes one after the other.  Consecutive
// passes that neither write one another's attachments nor sample them form
// a group; inside a group, the k-th launch of every pass (its k-th run of
// draws that share a shader program, split_pass) is merged with those of
// the other passes that use the same program into one launch with one
// render target per pass.
int plan_passes(const cpv_pass_desc* passes, uint32_t pass_count, const cpv_pipeline_desc* pipelines, const cpv_draw_desc* draws,
                const uint32_t* draw_pipeline, const uint32_t* first_draw, int tile, std::vector<LaunchSpec>* out)
{
    uint32_t a = 0;
    while (a < pass_count) {
        // the group [a, b): each pass independent of the ones before it in the group
        std::vector<Span> written, sampled;
        uint32_t b = a;
        for (; b < pass_count; b++) {
            std::vector<Span> w, r;
            attachment_spans(passes[b], &w);
            for (uint32_t k = first_draw[b]; k < first_draw[b + 1]; k++) texture_spans(draws[k], &r);
            if (b > a && (spans_meet(w, written) || spans_meet(w, sampled) || spans_meet(r, written))) break;
            written.insert(written.end(), w.begin(), w.end());
            sampled.insert(sampled.end(), r.begin(), r.end());
        }
        std::vector<std::vector<PassLaunch>> runs(b - a);
        size_t steps = 0;
        for (uint32_t p = a; p < b; p++) {
            const uint32_t f = first_draw[p], n = first_draw[p + 1] - f;
            int rc = split_pass(pipelines, draws + f, draw_pipeline ? draw_pipeline + f : nullptr, n, &runs[p - a]);
            if (rc) return rc;
            steps = std::max(steps, runs[p - a].size());
        }
        for (size_t k = 0; k < steps; k++) {
            std::vector<bool> taken(b - a, false);
            for (uint32_t p = a; p < b; p++) {
                if (taken[p - a] || runs[p - a].size() <= k) continue;
                // runs without draws (clear-only passes) join any launch
                const PassLaunch* by = nullptr;   // the run whose program the launch runs
                LaunchSpec G;
                uint32_t tile_base = 0;
                for (uint32_t q = p; q < b; q++) {
                    if (taken[q - a] || runs[q - a].size() <= k) continue;
                    const PassLaunch& R = runs[q - a][k];
                    if (!R.lds.empty()) {
                        if (by && !same_program(*R.lds[0].pipe, *by->lds[0].pipe)) continue;
                        if (!by) by = &R;
                    }
                    taken[q - a] = true;
                    TargetParams t;
                    int rc = build_target(passes[q], R.first, tile, tile_base, &t);
                    if (rc) return rc;
                    if ((uint64_t)tile_base + (uint64_t)t.tiles_x * t.tiles_y >= (1ull << 31))
                        return fail(CPV_ERROR_UNSUPPORTED, "more than 2^31 tiles in one launch");
                    tile_base += t.tiles_x * t.tiles_y;
                    const uint32_t ti = (uint32_t)G.targets.size();
                    G.targets.push_back(t);
                    G.passes.push_back(passes[q]);
                    for (const LaunchDraw& d : R.lds) {
                        G.lds.push_back(d);
                        G.dt.push_back(ti);
                    }
                }
                G.prog = by ? by->prog : runs[p - a][k].prog;
                out->push_back(std::move(G));
            }
        }
        a = b;
    }
    return CPV_OK;
}

int check_passes_args(cpv_context* c, const cpv_pass_desc* passes, uint32_t pass_count, const cpv_pipeline_desc* pipelines,
                      const cpv_draw_desc* draws, const uint32_t* draw_pipeline, const uint32_t* first_draw)
{
    if (!c || (pass_count && (!passes || !first_draw))) return fail(CPV_ERROR_INVALID, "null argument");
    if (pass_count && first_draw[0] != 0) return fail(CPV_ERROR_INVALID, "first_draw[0] is %u, not 0", first_draw[0]);
    for (uint32_t p = 0; p < pass_count; p++) {
        if (first_draw[p + 1] < first_draw[p]) return fail(CPV_ERROR_INVALID, "first_draw decreases at pass %u", p + 1);
        if (passes[p].color_count > CPV_MAX_COLOR_ATTACHMENTS)
            return fail(CPV_ERROR_UNSUPPORTED, "%u colour attachments", passes[p].color_count);
    }
    const uint32_t n = pass_count ? first_draw[pass_count] : 0;
    if (n && (!pipelines || !draws)) return fail(CPV_ERROR_INVALID, "null argument");
    (void)draw_pipeline;
    return CPV_OK;
}

}  // namespace

extern "C" int cpv_render_pass(cpv_context* c, const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                               const cpv_draw_desc* draws, const uint32_t* draw_pipeline, uint32_t draw_count)
{
    if (int rc = check_pass_args(c, pass, pipelines, draws, draw_count)) return rc;
    std::lock_guard<std::mutex> guard(c->lock);
    if (int rc = begin_render_call(c)) return rc;
    const cpv_config cfg = call_config(c->cfg, pass_has_points_or_lines(pipelines, draw_pipeline, draw_count));
    const std::vector<cpv_pass_desc> passes(1, *pass);
    std::vector<PassLaunch> launches;
    int rc = split_pass(pipelines, draws, draw_pipeline, draw_count, &launches);
    if (rc) return rc;
    for (const PassLaunch& L : launches) {
        std::vector<TargetParams> targets(1);
        rc = build_target(*pass, L.first, (int)cfg.tile_size, 0, &targets[0]);
        if (rc) return rc;
        rc = run_launch(c, cfg, L.prog, targets, L.lds, passes, L.dt);
        if (rc) return rc;
    }
    return end_render_call(c);
}

// cpv_render_pass with every working buffer in this GPU's scratch array:
// every launch of the pass is planned and checked against the array and the
// staging slot first, so a pass that does not fit is refused before any of
// it is enqueued; then the launches run one after the other in the array.
// No hipMalloc, hipFree, hipHostMalloc or other allocating call is made.
extern "C" int cpv_render_pass_noalloc(cpv_context* c, const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                                       const cpv_draw_desc* draws, const uint32_t* draw_pipeline, uint32_t draw_count)
{
    if (int rc = check_pass_args(c, pass, pipelines, draws, draw_count)) return rc;
    std::lock_guard<std::mutex> guard(c->lock);
    if (!c->scratch) {
        if (int rc = scratch_setup(c)) return rc;   // reports why the context could not take the array's address
    }
    HIP_TRY(hipSetDevice(c->device));
    cpv_config cfg = call_config(c->cfg, pass_has_points_or_lines(pipelines, draw_pipeline, draw_count));
    cfg.record_costs = 0;
    const std::vector<cpv_pass_desc> passes(1, *pass);
    std::vector<PassLaunch> split;
    int rc = split_pass(pipelines, draws, draw_pipeline, draw_count, &split);
    if (rc) return rc;
    std::vector<std::unique_ptr<PendingLaunch>> launches;
    for (const PassLaunch& S : split) {
        std::vector<TargetParams> targets(1);
        rc = build_target(*pass, S.first, (int)cfg.tile_size, 0, &targets[0]);
        if (rc) return rc;
        std::unique_ptr<PendingLaunch> L(new PendingLaunch());
        L->cfg = cfg;
        rc = plan_launch(cfg, S.prog, targets, S.lds, passes, S.dt, &L->P);
        if (rc) return rc;
        stage_plan(L->P, L->st);
        if (L->st.bytes.size() > kStagingSlotBytes)
            return fail(CPV_ERROR_UNSUPPORTED, "the pass's descriptors take %zu bytes; the allocation-free entry point stages at most %zu",
                        L->st.bytes.size(), kStagingSlotBytes);
        Buffers probe;
        uint32_t cap = 0;
        rc = carve_scratch(c->scratch, L->P, L->st.bytes.size(), &probe, &cap);
        if (rc) return rc;
        launches.push_back(std::move(L));
    }
    note_enqueue(c);
    rc = resolve_pending(c, false);
    if (rc) return rc;
    fill_u32(c->stream, c->scratch, 0u, kStatWords);   // fragment counter, pair counts, overflow word
    HIP_TRY(hipEventRecord(c->ev_begin, c->stream));
    for (std::unique_ptr<PendingLaunch>& L : launches) {
        rc = resolve_pending(c, false);   // the previous launch of the pass, before this one reuses the array
        if (rc) return rc;
        rc = enqueue_scratch(c, *L);
        if (rc) return rc;
        c->last_tiles = L->P.tiles;
        c->last_waves = cfg.waves_per_tile;
        c->last_tile_size = cfg.tile_size;
        c->last_traces = false;
        c->last_target_base.assign(1, 0u);
        c->pending = L.release();
        c->launches++;
    }
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    c->timing_valid = true;
    c->last_count_exact = cfg.count_fragments != 0;
    c->last_stats = (uint32_t*)c->scratch;
    return CPV_OK;
}

extern "C" int cpv_render_jobs(cpv_context* c, const cpv_job_desc* jobs, uint32_t job_count)
{
    if (!c || (job_count && !jobs)) return fail(CPV_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> guard(c->lock);
    if (int rc = begin_render_call(c)) return rc;
    const cpv_config cfg = call_config(c->cfg, jobs_have_points_or_lines(jobs, job_count));
    std::vector<LaunchSpec> groups;
    int rc = group_jobs(jobs, job_count, (int)cfg.tile_size, &groups);
    if (rc) return rc;
    for (LaunchSpec& G : groups) {
        rc = run_launch(c, cfg, G.prog, G.targets, G.lds, G.passes, G.dt);
        if (rc) return rc;
    }
    return end_render_call(c);
}

namespace {

bool passes_have_points_or_lines(const cpv_pipeline_desc* pipelines, const uint32_t* draw_pipeline, uint32_t draw_count)
{
    return draw_count && pass_has_points_or_lines(pipelines, draw_pipeline, draw_count);
}

}  // namespace

extern "C" int cpv_render_passes(cpv_context* c, const cpv_pass_desc* passes, uint32_t pass_count, const cpv_pipeline_desc* pipelines,
                                 const cpv_draw_desc* draws, const uint32_t* draw_pipeline, const uint32_t* first_draw)
{
    if (int rc = check_passes_args(c, passes, pass_count, pipelines, draws, draw_pipeline, first_draw)) return rc;
    std::lock_guard<std::mutex> guard(c->lock);
    if (int rc = begin_render_call(c)) return rc;
    const uint32_t n = pass_count ? first_draw[pass_count] : 0;
    const cpv_config cfg = call_config(c->cfg, passes_have_points_or_lines(pipelines, draw_pipeline, n));
    std::vector<LaunchSpec> specs;
    int rc = plan_passes(passes, pass_count, pipelines, draws, draw_pipeline, first_draw, (int)cfg.tile_size, &specs);
    if (rc) return rc;
    for (LaunchSpec& G : specs) {
        rc = run_launch(c, cfg, G.prog, G.targets, G.lds, G.passes, G.dt);
        if (rc) return rc;
    }
    return end_render_call(c);
}

extern "C" int cpv_last_fragment_count(cpv_context* c, uint64_t* fragments)
{
    if (!c || !fragments) return fail(CPV_ERROR_INVALID, "null argument");
    if (!c->last_count_exact) return fail(CPV_ERROR_INVALID, "the last render call ran with count_fragments off");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipMemcpyAsync(c->readback + 16, c->last_stats, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(fragments, c->readback + 16, 8);
    return CPV_OK;
}

extern "C" int cpv_last_timing(cpv_context* c, float* raster_ms, float* total_ms)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    if (!c->timing_valid) return fail(CPV_ERROR_INVALID, "no render call recorded");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipEventSynchronize(c->ev_end));
    if (raster_ms) HIP_TRY(hipEventElapsedTime(raster_ms, c->ev_raster0, c->ev_raster1));
    if (total_ms) HIP_TRY(hipEventElapsedTime(total_ms, c->ev_begin, c->ev_end));
    return CPV_OK;
}

extern "C" int cpv_get_tile_traces(cpv_context* c, cpv_tile_trace* traces, uint32_t capacity, uint32_t* count)
{
    if (!c || !count || (capacity && !traces)) return fail(CPV_ERROR_INVALID, "null argument");
    *count = 0;
    if (!c->last_traces) return fail(CPV_ERROR_INVALID, "the last launch ran with record_costs off");
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    const uint32_t tiles = c->last_tiles, waves = c->last_waves;
    *count = tiles;
    const uint32_t n = std::min(tiles, capacity);
    if (n == 0) return CPV_OK;
    std::vector<float> feat((size_t)n * kFeatureStride), cyc((size_t)n * waves);
    HIP_TRY(hipStreamSynchronize(c->stream));
    const Buffers& B = c->lbuf[c->last_set];
    HIP_TRY(hipMemcpy(feat.data(), B.features.p, feat.size() * 4, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cyc.data(), B.traces.p, cyc.size() * 4, hipMemcpyDeviceToHost));
    for (uint32_t t = 0; t < n; t++) {
        cpv_tile_trace& o = traces[t];
        std::memset(&o, 0, sizeof(o));
        for (int k = 0; k < kFeatures; k++) o.features[k] = feat[(size_t)t * kFeatureStride + k];
        float s = 0.0f;
        for (uint32_t w = 0; w < waves; w++) s += cyc[(size_t)t * waves + w];
        o.cycles = s;
        const auto it = std::upper_bound(c->last_target_base.begin(), c->last_target_base.end(), t);
        const uint32_t job = (uint32_t)(it - c->last_target_base.begin()) - 1u;
        o.job = job;
        o.tile = t - c->last_target_base[job];
    }
    return CPV_OK;
}

// -------------------------------------------------------- recorded batches

// A recorded batch holds two sets of working buffers, and each set two
// graphs: the front graph (vertex stage, setup, binning, scheduling; no
// attachment is touched) on the context's front stream, and the raster
// graph (the raster kernels) on the context's stream.  Submit n uses set
// n % 2.  Its raster graph waits for its front graph and, by stream order,
// for the previous submit's raster graph; its front graph waits only for the
// raster graph that last used the same set.  So when submits follow each
// other with nothing else enqueued on the context in between, the front
// half of one frame runs while the raster kernel of the previous frame
// does.  Any other work enqueued on the context in between makes the next
// front graph wait for all of it.
//
// The launches of a set share one overflow word, cleared at the start of
// its front graph and raised by any launch whose pairs exceed the capacity
// recorded for it; every raster kernel of the set then draws nothing, so a
// submit either draws all of its frame or none of it.  The word is read
// back after the front graph, and the next call on the context checks it
// (resolve_pending): a submit that drew nothing is recorded again with the
// pair counts it needed and replayed before anything else is enqueued.
struct cpv_batch {
    struct Set {
        std::vector<cpvrt::Buffers> b;   // one per launch of the frame
        hipGraphExec_t front = nullptr, raster = nullptr;
        hipEvent_t front_done = nullptr, raster_done = nullptr;
        bool used = false;
        uint32_t* overflow() const { return b[0].stats.as<uint32_t>() + kStatOverflow; }
    };
    cpv_context* c = nullptr;
    cpv_config cfg{};
    cpvrt::DevBuf arena;
    std::vector<LaunchPlan> plans;
    std::vector<uint32_t> caps;
    std::vector<cpvrt::DevBuf> order;   // per launch: the learned tile order (PREDICTED), computed once
    Set set[2];
    hipEvent_t e0 = nullptr, e1 = nullptr, pre = nullptr;
    uint64_t submits = 0;
    uint64_t last_seq = 0;   // the context's enqueue count after this batch's last submit
    bool submitted = false;
};

static void release_graphs(cpv_batch* B)
{
    for (cpv_batch::Set& S : B->set) {
        if (S.front) (void)hipGraphExecDestroy(S.front);
        if (S.raster) (void)hipGraphExecDestroy(S.raster);
        S.front = S.raster = nullptr;
    }
}

// Frees a batch; the caller holds the context's lock and has resolved any
// pending check of the batch.
static void free_batch(cpv_batch* B)
{
    (void)hipSetDevice(B->c->device);
    (void)sync_context(B->c);
    release_graphs(B);
    for (cpv_batch::Set& S : B->set) {
        if (S.front_done) (void)hipEventDestroy(S.front_done);
        if (S.raster_done) (void)hipEventDestroy(S.raster_done);
        for (cpvrt::Buffers& b : S.b) b.release_all();
    }
    for (cpvrt::DevBuf& o : B->order) o.release();
    for (hipEvent_t e : {B->e0, B->e1, B->pre})
        if (e) (void)hipEventDestroy(e);
    B->arena.release();
    delete B;
}

extern "C" void cpv_batch_destroy(cpv_batch* B)
{
    if (!B) return;
    std::lock_guard<std::mutex> guard(B->c->lock);
    if (B->c->pending_batch == B) (void)resolve_pending(B->c);   // a submit that drew nothing is redrawn first
    free_batch(B);
}

// Captures what `body` enqueues on stream s into an instantiated graph.
template <class F>
static int capture(hipStream_t s, hipGraphExec_t* out, F&& body)
{
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    const int rc = body();
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(CPV_ERROR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(CPV_ERROR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    return CPV_OK;
}

// Sizes both sets' buffers for the batch's capacities, clears their control
// words, and records the four graphs: no allocation, no host sync and no
// memset node inside a capture.
// A submit's front half of set S on the context's front stream (what the
// set's front graph holds): every launch's vertex stage, setup and binning,
// into the set's buffers, with the launches' shared overflow word cleared
// first.  bins receives where each launch's binning leaves its results.
static int batch_front(cpv_context* c, cpv_batch* B, cpv_batch::Set& S, std::vector<Binned>& bins)
{
    uint8_t* arena = B->arena.as<uint8_t>();
    fill_u32(c->front_stream, S.overflow(), 0u, 1);
    for (size_t g = 0; g < B->plans.size(); g++) {
        Exec x = context_exec(c, &S.b[g]);
        x.cfg = B->cfg;
        x.s = c->front_stream;
        x.frozen = true;
        x.overflow = S.overflow();
        if (!B->order.empty() && B->order[g].p) x.fixed_order = B->order[g].as<uint32_t>();
        fill_u32(x.s, S.b[g].stats.p, 0u, 2);   // the fragment counter restarts
        int r = enqueue_front(x, B->plans[g], arena);
        if (r == CPV_OK) r = enqueue_bin(x, B->plans[g], arena, B->caps[g], &bins[g]);
        if (r) return r;
    }
    return CPV_OK;
}

// The raster kernels of a submit of set S on the context's stream (what the
// set's raster graph holds).
static int batch_raster(cpv_context* c, cpv_batch* B, cpv_batch::Set& S, const std::vector<Binned>& bins)
{
    uint8_t* arena = B->arena.as<uint8_t>();
    for (size_t g = 0; g < B->plans.size(); g++) {
        Exec x = context_exec(c, &S.b[g]);
        x.cfg = B->cfg;
        x.frozen = true;
        x.overflow = S.overflow();
        const int r = enqueue_raster(x, B->plans[g], arena, bins[g]);
        if (r) return r;
    }
    return CPV_OK;
}

static int batch_record(cpv_context* c, cpv_batch* B)
{
    const size_t G = B->plans.size();
    uint8_t* arena = B->arena.as<uint8_t>();
    for (cpv_batch::Set& S : B->set) {
        for (size_t g = 0; g < G; g++) {
            Exec y = context_exec(c, &S.b[g]);
            y.cfg = B->cfg;
            int rc = ensure_front(y, B->plans[g]);
            if (rc == CPV_OK) rc = ensure_back(y, B->plans[g], B->caps[g]);
            if (rc) return rc;
            if (S.b[g].stats.ensure(kStatAlloc * 4)) return CPV_ERROR_OUT_OF_MEMORY;
            fill_u32(c->stream, S.b[g].stats.p, 0u, kStatAlloc);
        }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    // The learned order: the front half and binning of every launch run
    // once, eagerly, on the first set's buffers, and the order they give is
    // kept for every submit.  It is a schedule only (tiles are disjoint), so
    // vertex data that changes between submits leaves the frames exact.
    Exec probe = context_exec(c, nullptr);
    probe.cfg = B->cfg;
    if (effective_scheduler(probe) == CPV_SCHED_PREDICTED && B->order.empty()) {
        B->order.resize(G);
        for (size_t g = 0; g < G; g++) {
            Exec x = context_exec(c, &B->set[0].b[g]);
            x.cfg = B->cfg;
            Binned bn;
            int rc = enqueue_front(x, B->plans[g], arena);
            if (rc == CPV_OK) rc = enqueue_bin(x, B->plans[g], arena, B->caps[g], &bn);
            if (rc) return rc;
            if (bn.order) {
                if (B->order[g].ensure((size_t)B->plans[g].tiles * 4)) return CPV_ERROR_OUT_OF_MEMORY;
                HIP_TRY(hipMemcpyAsync(B->order[g].p, bn.order, (size_t)B->plans[g].tiles * 4, hipMemcpyDeviceToDevice, c->stream));
            }
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (cpv_batch::Set& S : B->set)
            for (size_t g = 0; g < G; g++) fill_u32(c->stream, S.b[g].stats.p, 0u, kStatAlloc);
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    for (cpv_batch::Set& S : B->set) {
        std::vector<Binned> bins(G);
        int rc = capture(c->front_stream, &S.front, [&]() { return batch_front(c, B, S, bins); });
        if (rc) return rc;
        rc = capture(c->stream, &S.raster, [&]() { return batch_raster(c, B, S, bins); });
        if (rc) return rc;
    }
    return CPV_OK;
}

static int batch_build(cpv_context* c, cpv_batch* B, std::vector<LaunchSpec>& specs)
{
    for (hipEvent_t* e : {&B->e0, &B->e1}) HIP_TRY(hipEventCreate(e));
    HIP_TRY(hipEventCreateWithFlags(&B->pre, hipEventDisableTiming));
    if (!c->front_stream) HIP_TRY(hipStreamCreateWithFlags(&c->front_stream, hipStreamNonBlocking));
    const size_t G = specs.size();
    if (G == 0) return fail(CPV_ERROR_INVALID, "a batch needs at least one launch");
    B->plans.resize(G);
    Staging st;
    for (size_t g = 0; g < G; g++) {
        int rc = plan_launch(B->cfg, specs[g].prog, specs[g].targets, specs[g].lds, specs[g].passes, specs[g].dt, &B->plans[g]);
        if (rc) return rc;
        stage_plan(B->plans[g], st);
    }
    if (B->arena.ensure(std::max<size_t>(st.bytes.size(), 256))) return CPV_ERROR_OUT_OF_MEMORY;
    uint8_t* arena = B->arena.as<uint8_t>();
    for (LaunchPlan& P : B->plans) patch_plan(P, st, arena);
    int rc = upload_staging(c, B->arena, st);
    if (rc) return rc;
    // sizing run: the front half of every launch, to learn its pair count
    B->caps.resize(G);
    for (cpv_batch::Set& S : B->set) {
        S.b.resize(G);
        HIP_TRY(hipEventCreateWithFlags(&S.front_done, hipEventDisableTiming));
        HIP_TRY(hipEventCreateWithFlags(&S.raster_done, hipEventDisableTiming));
    }
    for (size_t g = 0; g < G; g++) {
        Exec x = context_exec(c, &B->set[0].b[g]);
        x.cfg = B->cfg;
        rc = enqueue_front(x, B->plans[g], arena);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(c->readback, x.b->stats.as<uint32_t>() + kStatPairTotal, 8, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        uint64_t pairs = 0;
        std::memcpy(&pairs, c->readback, 8);
        if (pairs > (uint64_t)UINT32_MAX)
            return fail(CPV_ERROR_OVERFLOW, "the batch's launch %zu needs %llu (tile, primitive) pairs, more than 2^32 - 1", g,
                        (unsigned long long)pairs);
        // room for vertex data that moves a little between submits
        B->caps[g] = (uint32_t)std::min<uint64_t>(pairs + pairs / 8 + 4096, UINT32_MAX);
        if (resident_workgroups(x, B->plans[g].prog.raster[B->plans[g].raster_variant], 64 * (int)B->cfg.waves_per_tile) <= 0)
            return CPV_ERROR_DEVICE;
    }
    return batch_record(c, B);
}

// Enqueues one replay of set si (front graph, read-back of its overflow word
// for the next call's check, raster graph).
static int batch_launch(cpv_context* c, cpv_batch* B, uint32_t si, bool chained)
{
    cpv_batch::Set& S = B->set[si];
    if (!chained) {
        // other work was enqueued since this batch's last submit: the front
        // half waits for all of it
        HIP_TRY(hipEventRecord(B->pre, c->stream));
        HIP_TRY(hipStreamWaitEvent(c->front_stream, B->pre, 0));
    }
    if (S.used) HIP_TRY(hipStreamWaitEvent(c->front_stream, S.raster_done, 0));   // the set's buffers are free again
    HIP_TRY(hipEventRecord(B->e0, c->front_stream));
    HIP_TRY(hipGraphLaunch(S.front, c->front_stream));
    HIP_TRY(hipMemcpyAsync(c->readback + 36, S.overflow(), 4, hipMemcpyDeviceToHost, c->front_stream));
    HIP_TRY(hipEventRecord(c->ev_check, c->front_stream));
    HIP_TRY(hipEventRecord(S.front_done, c->front_stream));
    HIP_TRY(hipStreamWaitEvent(c->stream, S.front_done, 0));
    HIP_TRY(hipGraphLaunch(S.raster, c->stream));
    HIP_TRY(hipEventRecord(S.raster_done, c->stream));
    HIP_TRY(hipEventRecord(B->e1, c->stream));
    S.used = true;
    c->pending_batch = B;
    c->pending_set = si;
    c->launches += B->plans.size();
    return CPV_OK;
}

// A submit of set si raised its overflow word and drew nothing: every
// launch of the batch is given the capacity its pairs needed (x 1.125 +
// 4096), both sets are recorded again, and the submit is replayed.
int cpvrt::batch_redo(cpv_context* c, cpv_batch* B, uint32_t si, bool may_allocate)
{
    HIP_TRY(sync_context(c));
    cpv_batch::Set& S = B->set[si];
    for (size_t g = 0; g < B->plans.size(); g++) {
        uint64_t total = 0;
        HIP_TRY(hipMemcpy(&total, S.b[g].stats.as<uint32_t>() + kStatPairTotal, 8, hipMemcpyDeviceToHost));
        if (total > (uint64_t)UINT32_MAX)
            return fail(CPV_ERROR_OVERFLOW,
                        "a submit of a recorded batch needs %llu (tile, primitive) pairs in launch %zu, more than 2^32 - 1 "
                        "(nothing was drawn)",
                        (unsigned long long)total, g);
        if (total > B->caps[g]) {
            if (!may_allocate)
                return fail(CPV_ERROR_OVERFLOW, "a submit of a recorded batch outgrew its pair arrays (nothing was drawn); "
                                                "call cpv_synchronize before allocation-free calls");
            B->caps[g] = (uint32_t)std::min<uint64_t>(total + total / 8 + 4096, UINT32_MAX);
        }
    }
    release_graphs(B);
    int rc = batch_record(c, B);
    if (rc) return rc;
    B->set[0].used = B->set[1].used = false;
    return batch_launch(c, B, si, false);
}

static int create_batch(cpv_context* c, const cpv_config& cfg, std::vector<LaunchSpec>& specs, cpv_batch** out)
{
    cpv_batch* B = new cpv_batch();
    B->c = c;
    B->cfg = cfg;
    int rc = batch_build(c, B, specs);
    if (rc) {
        const std::string msg = cpv_last_error();
        free_batch(B);
        return fail(rc, "%s", msg.c_str());
    }
    note_enqueue(c);
    *out = B;
    return CPV_OK;
}

extern "C" int cpv_batch_create(cpv_context* c, const cpv_job_desc* jobs, uint32_t job_count, cpv_batch** out)
{
    if (!c || !out || (job_count && !jobs)) return fail(CPV_ERROR_INVALID, "null argument");
    if (job_count == 0) return fail(CPV_ERROR_INVALID, "a batch needs at least one job");
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    const cpv_config cfg = call_config(c->cfg, jobs_have_points_or_lines(jobs, job_count));
    std::vector<LaunchSpec> specs;
    if (int rc = group_jobs(jobs, job_count, (int)cfg.tile_size, &specs)) return rc;
    return create_batch(c, cfg, specs, out);
}

extern "C" int cpv_batch_create_passes(cpv_context* c, const cpv_pass_desc* passes, uint32_t pass_count, const cpv_pipeline_desc* pipelines,
                                       const cpv_draw_desc* draws, const uint32_t* draw_pipeline, const uint32_t* first_draw,
                                       cpv_batch** out)
{
    if (!out) return fail(CPV_ERROR_INVALID, "null argument");
    if (int rc = check_passes_args(c, passes, pass_count, pipelines, draws, draw_pipeline, first_draw)) return rc;
    if (pass_count == 0) return fail(CPV_ERROR_INVALID, "a batch needs at least one render pass");
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    const cpv_config cfg = call_config(c->cfg, passes_have_points_or_lines(pipelines, draw_pipeline, first_draw[pass_count]));
    std::vector<LaunchSpec> specs;
    if (int rc = plan_passes(passes, pass_count, pipelines, draws, draw_pipeline, first_draw, (int)cfg.tile_size, &specs)) return rc;
    return create_batch(c, cfg, specs, out);
}

extern "C" int cpv_batch_submit(cpv_batch* B)
{
    if (!B) return fail(CPV_ERROR_INVALID, "null batch");
    cpv_context* c = B->c;
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    const bool chained = B->submits > 0 && B->last_seq == c->enqueue_seq;
    int rc = batch_launch(c, B, (uint32_t)(B->submits & 1u), chained);
    if (rc) return rc;
    B->submits++;
    B->submitted = true;
    note_enqueue(c);
    B->last_seq = c->enqueue_seq;
    return CPV_OK;
}

extern "C" int cpv_batch_wait(cpv_batch* B, uint64_t* fragments)
{
    if (!B) return fail(CPV_ERROR_INVALID, "null batch");
    if (!B->submitted) return fail(CPV_ERROR_INVALID, "batch not submitted");
    cpv_context* c = B->c;
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipEventSynchronize(B->e1));
    uint64_t frags = 0;
    const cpv_batch::Set& last = B->set[(B->submits - 1) & 1u];
    for (const cpvrt::Buffers& b : last.b) {
        uint64_t f = 0;
        HIP_TRY(hipMemcpy(&f, b.stats.as<uint32_t>() + kStatFragments, 8, hipMemcpyDeviceToHost));
        frags += f;
    }
    if (fragments) {
        if (!B->cfg.count_fragments) return fail(CPV_ERROR_INVALID, "the batch was recorded with count_fragments off");
        *fragments = frags;
    }
    return CPV_OK;
}

extern "C" int cpv_batch_timing(cpv_batch* B, float* ms)
{
    if (!B || !ms) return fail(CPV_ERROR_INVALID, "null argument");
    if (!B->submitted) return fail(CPV_ERROR_INVALID, "batch not submitted");
    std::lock_guard<std::mutex> guard(B->c->lock);
    HIP_TRY(hipSetDevice(B->c->device));
    if (int rc = resolve_pending(B->c)) return rc;
    HIP_TRY(hipEventSynchronize(B->e1));
    HIP_TRY(hipEventElapsedTime(ms, B->e0, B->e1));
    return CPV_OK;
}

// ------------------------------------------------------------------ fences

// A fence marks the end of the work enqueued on a context up to its signal
// (Queue::Submit signalling its VkFence after the command buffers,
// CPVulkan/Queue.cpp:11-77; Fence::Wait / getStatus, CPVulkan/Fence.cpp:
// 47-56, 157-160).  It is a HIP event recorded on the context's stream,
// which every launch and transfer ends on.  A fence signalled after a
// launch whose pair count is still unchecked is recorded again after the
// redraw if that launch has to be redrawn.
struct cpv_fence {
    cpv_context* c = nullptr;
    hipEvent_t ev = nullptr;
    bool armed = false;      // an event was recorded by cpv_fence_signal
    bool signaled = false;   // created signalled, with no signal since
};

static void unlist_fence(cpv_context* c, cpv_fence* f)
{
    auto& v = c->fences_after_pending;
    v.erase(std::remove(v.begin(), v.end(), f), v.end());
}

void cpvrt::rerecord_fences(cpv_context* c)
{
    for (cpv_fence* f : c->fences_after_pending) (void)hipEventRecord(f->ev, c->stream);
}

extern "C" int cpv_fence_create(cpv_context* c, int signaled, cpv_fence** out)
{
    if (!c || !out) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    cpv_fence* f = new cpv_fence();
    f->c = c;
    f->signaled = signaled != 0;
    if (hipEventCreateWithFlags(&f->ev, hipEventDisableTiming) != hipSuccess) {
        delete f;
        return fail(CPV_ERROR_DEVICE, "hipEventCreate failed");
    }
    *out = f;
    return CPV_OK;
}

extern "C" int cpv_fence_signal(cpv_fence* f)
{
    if (!f) return fail(CPV_ERROR_INVALID, "null fence");
    cpv_context* c = f->c;
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventRecord(f->ev, c->stream));
    f->armed = true;
    f->signaled = false;
    unlist_fence(c, f);
    if (c->pending || c->pending_batch) c->fences_after_pending.push_back(f);
    return CPV_OK;
}

extern "C" int cpv_fence_reset(cpv_fence* f)
{
    if (!f) return fail(CPV_ERROR_INVALID, "null fence");
    std::lock_guard<std::mutex> guard(f->c->lock);
    unlist_fence(f->c, f);
    f->armed = false;
    f->signaled = false;
    return CPV_OK;
}

extern "C" int cpv_fence_status(cpv_fence* f)
{
    if (!f) return fail(CPV_ERROR_INVALID, "null fence");
    cpv_context* c = f->c;
    std::lock_guard<std::mutex> guard(c->lock);
    if (!f->armed) return f->signaled ? CPV_OK : CPV_NOT_READY;
    HIP_TRY(hipSetDevice(c->device));
    for (int pass = 0; pass < 2; pass++) {
        const hipError_t e = hipEventQuery(f->ev);
        if (e == hipErrorNotReady) return CPV_NOT_READY;
        if (e != hipSuccess) return fail(CPV_ERROR_DEVICE, "hipEventQuery: %s", hipGetErrorString(e));
        const auto& v = c->fences_after_pending;
        if (std::find(v.begin(), v.end(), f) == v.end()) return CPV_OK;
        // the work before the fence is done, and with it the pair-count
        // check it depends on: resolve it (a redraw records the fence again)
        if (int rc = resolve_pending(c)) return rc;
    }
    return CPV_NOT_READY;
}

extern "C" int cpv_fence_wait(cpv_fence* f, uint64_t timeout_ns)
{
    if (!f) return fail(CPV_ERROR_INVALID, "null fence");
    const auto start = std::chrono::steady_clock::now();
    for (;;) {
        const int st = cpv_fence_status(f);
        if (st != CPV_NOT_READY) return st;
        const uint64_t waited =
            (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - start).count();
        if (waited >= timeout_ns) return CPV_TIMEOUT;
        if (f->armed && timeout_ns == UINT64_MAX) {
            (void)hipEventSynchronize(f->ev);
        } else {
            const uint64_t left = timeout_ns - waited;
            std::this_thread::sleep_for(std::chrono::nanoseconds(std::min<uint64_t>(left, 20000)));
        }
    }
}

extern "C" void cpv_fence_destroy(cpv_fence* f)
{
    if (!f) return;
    {
        std::lock_guard<std::mutex> guard(f->c->lock);
        unlist_fence(f->c, f);
    }
    (void)hipEventDestroy(f->ev);
    delete f;
}

extern "C" int cpv_launch_count(cpv_context* c, uint64_t* launches)
{
    if (!c || !launches) return fail(CPV_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> guard(c->lock);
    *launches = c->launches;
    return CPV_OK;
}

// ---------------------------------------------------------------- autotuner

extern "C" int cpv_autotune(cpv_context* c, const cpv_job_desc* jobs, uint32_t job_count, uint32_t repeats,
                            cpv_autotune_entry* entries, uint32_t capacity, uint32_t* count, cpv_config* best)
{
    if (!c || !count || (capacity && !entries)) return fail(CPV_ERROR_INVALID, "null argument");
    static const uint32_t configs[7][2] = {{8, 1}, {16, 1}, {16, 2}, {16, 4}, {32, 1}, {32, 2}, {32, 4}};
    const cpv_config saved = c->cfg;
    cpv_config best_cfg = saved;
    float best_ms = -1.0f;
    uint32_t n = 0;
    repeats = std::max(repeats, 1u);
    for (const auto& tc : configs) {
        cpv_config cfg = saved;
        cfg.tile_size = tc[0];
        cfg.waves_per_tile = tc[1];
        cfg.record_costs = 0;
        c->cfg = cfg;
        cpv_batch* B = nullptr;
        int rc = cpv_batch_create(c, jobs, job_count, &B);
        if (rc) { c->cfg = saved; return rc; }
        float total = 0.0f;
        // one untimed replay, then `repeats` timed ones
        for (uint32_t r = 0; r <= repeats && rc == CPV_OK; r++) {
            rc = cpv_batch_submit(B);
            if (rc == CPV_OK) rc = cpv_batch_wait(B, nullptr);
            float ms = 0.0f;
            if (rc == CPV_OK) rc = cpv_batch_timing(B, &ms);
            if (r > 0) total += ms;
        }
        cpv_batch_destroy(B);
        if (rc) { c->cfg = saved; return rc; }
        const float ms = total / (float)repeats;
        if (n < capacity) entries[n] = cpv_autotune_entry{tc[0], tc[1], ms, 0u};
        n++;
        if (best_ms < 0.0f || ms < best_ms) { best_ms = ms; best_cfg = cfg; }
    }
    c->cfg = saved;
    best_cfg.record_costs = saved.record_costs;
    *count = n;
    if (best) *best = best_cfg;
    return CPV_OK;
}

extern "C" int cpv_set_predictor(cpv_context* c, cpv_mlp* m)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    std::lock_guard<std::mutex> guard(c->lock);
    c->predictor = m;
    return CPV_OK;
}
