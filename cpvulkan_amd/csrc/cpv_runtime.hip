// Host runtime of libcpv.so: the C-ABI declared in include/cpv.h.
//
// A render pass becomes, per run of draws that share a shader program, one
// launch of each stage on the context's stream:
//   vertex kernel -> setup kernel -> scan -> emit -> tile scan ->
//   radix sort by tile (2-3 passes) -> raster kernel
// The raster kernel fuses the render pass clear (loadOp CLEAR) into its
// tile initialisation, so a frame touches each framebuffer byte once on the
// way in (LOAD only) and once on the way out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpv.h"
#include "bin_kernels.h"
#include "cpv_program.h"
#include "cpv_types.h"
#include "programs_builtin.h"
#include "raster_kernels.h"

using namespace cpv;

// ------------------------------------------------------------------ errors

static thread_local std::string g_error;

static int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                            \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) return fail(CPV_ERROR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" const char* cpv_last_error(void) { return g_error.c_str(); }

// ----------------------------------------------------------- format table

// Normal and depth/stencil rows of the reference's format table
// (CPVulkanBase/Formats.cpp:209-444) for the formats the draw path handles.
bool cpv_format_desc(uint32_t f, FormatDesc* o)
{
    std::memset(o, 0, sizeof(*o));
    o->vk_format = f;
    for (int i = 0; i < 4; i++) o->off[i] = -1;
    o->depth_off = o->stencil_off = -1;
    int ch = 0, el = 1, base = kBaseNone;
    bool bgr = false;
    switch (f) {
    case CPV_FORMAT_R8_UNORM: ch = 1; base = kBaseUNorm; break;
    case CPV_FORMAT_R8_SNORM: ch = 1; base = kBaseSNorm; break;
    case CPV_FORMAT_R8_USCALED: ch = 1; base = kBaseUScaled; break;
    case CPV_FORMAT_R8_SSCALED: ch = 1; base = kBaseSScaled; break;
    case CPV_FORMAT_R8_UINT: ch = 1; base = kBaseUInt; break;
    case CPV_FORMAT_R8_SINT: ch = 1; base = kBaseSInt; break;
    case CPV_FORMAT_R8G8_UNORM: ch = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R8G8_SNORM: ch = 2; base = kBaseSNorm; break;
    case CPV_FORMAT_R8G8_UINT: ch = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R8G8_SINT: ch = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R8G8B8_UNORM: ch = 3; base = kBaseUNorm; break;
    case CPV_FORMAT_R8G8B8_SNORM: ch = 3; base = kBaseSNorm; break;
    case CPV_FORMAT_R8G8B8_UINT: ch = 3; base = kBaseUInt; break;
    case CPV_FORMAT_R8G8B8_SINT: ch = 3; base = kBaseSInt; break;
    case CPV_FORMAT_B8G8R8_UNORM: ch = 3; base = kBaseUNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8_SNORM: ch = 3; base = kBaseSNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8_UINT: ch = 3; base = kBaseUInt; bgr = true; break;
    case CPV_FORMAT_B8G8R8_SINT: ch = 3; base = kBaseSInt; bgr = true; break;
    case CPV_FORMAT_R8G8B8A8_UNORM: ch = 4; base = kBaseUNorm; break;
    case CPV_FORMAT_R8G8B8A8_SNORM: ch = 4; base = kBaseSNorm; break;
    case CPV_FORMAT_R8G8B8A8_USCALED: ch = 4; base = kBaseUScaled; break;
    case CPV_FORMAT_R8G8B8A8_SSCALED: ch = 4; base = kBaseSScaled; break;
    case CPV_FORMAT_R8G8B8A8_UINT: ch = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R8G8B8A8_SINT: ch = 4; base = kBaseSInt; break;
    case CPV_FORMAT_B8G8R8A8_UNORM: ch = 4; base = kBaseUNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8A8_SNORM: ch = 4; base = kBaseSNorm; bgr = true; break;
    case CPV_FORMAT_B8G8R8A8_UINT: ch = 4; base = kBaseUInt; bgr = true; break;
    case CPV_FORMAT_B8G8R8A8_SINT: ch = 4; base = kBaseSInt; bgr = true; break;
    case CPV_FORMAT_R16_UNORM: ch = 1; el = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R16_SNORM: ch = 1; el = 2; base = kBaseSNorm; break;
    case CPV_FORMAT_R16_UINT: ch = 1; el = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R16_SINT: ch = 1; el = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R16G16_UNORM: ch = 2; el = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R16G16_UINT: ch = 2; el = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R16G16_SINT: ch = 2; el = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R16G16B16A16_UNORM: ch = 4; el = 2; base = kBaseUNorm; break;
    case CPV_FORMAT_R16G16B16A16_SNORM: ch = 4; el = 2; base = kBaseSNorm; break;
    case CPV_FORMAT_R16G16B16A16_UINT: ch = 4; el = 2; base = kBaseUInt; break;
    case CPV_FORMAT_R16G16B16A16_SINT: ch = 4; el = 2; base = kBaseSInt; break;
    case CPV_FORMAT_R32_UINT: ch = 1; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32_SINT: ch = 1; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32_SFLOAT: ch = 1; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_R32G32_UINT: ch = 2; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32G32_SINT: ch = 2; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32G32_SFLOAT: ch = 2; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_R32G32B32_UINT: ch = 3; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32G32B32_SINT: ch = 3; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32G32B32_SFLOAT: ch = 3; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_R32G32B32A32_UINT: ch = 4; el = 4; base = kBaseUInt; break;
    case CPV_FORMAT_R32G32B32A32_SINT: ch = 4; el = 4; base = kBaseSInt; break;
    case CPV_FORMAT_R32G32B32A32_SFLOAT: ch = 4; el = 4; base = kBaseSFloat; break;
    case CPV_FORMAT_D16_UNORM: o->total = 2; o->element = 2; o->base = kBaseUNorm; o->depth_off = 0; o->depth_kind = 16; return true;
    case CPV_FORMAT_X8_D24_UNORM_PACK32: o->total = 4; o->element = 4; o->base = kBaseUNorm; o->depth_off = 0; o->depth_kind = 24; return true;
    case CPV_FORMAT_D32_SFLOAT: o->total = 4; o->element = 4; o->base = kBaseSFloat; o->depth_off = 0; o->depth_kind = 32; return true;
    case CPV_FORMAT_S8_UINT: o->total = 1; o->element = 1; o->base = kBaseUInt; o->stencil_off = 0; return true;
    case CPV_FORMAT_D16_UNORM_S8_UINT: o->total = 3; o->element = 2; o->base = kBaseUNorm; o->depth_off = 0; o->stencil_off = 2; o->depth_kind = 16; return true;
    case CPV_FORMAT_D24_UNORM_S8_UINT: o->total = 4; o->element = 3; o->base = kBaseUNorm; o->depth_off = 0; o->stencil_off = 3; o->depth_kind = 24; return true;
    case CPV_FORMAT_D32_SFLOAT_S8_UINT: o->total = 8; o->element = 4; o->base = kBaseSFloat; o->depth_off = 0; o->stencil_off = 4; o->depth_kind = 32; return true;
    default: return false;
    }
    o->base = base;
    o->element = el;
    o->total = el * ch;
    static const int rgba[4] = {0, 1, 2, 3}, bgra[4] = {2, 1, 0, 3};
    for (int i = 0; i < ch; i++) o->off[i] = (bgr ? bgra[i] : rgba[i]) * el;
    return true;
}

static bool is_depth_stencil(const FormatDesc& f) { return f.depth_off >= 0 || f.stencil_off >= 0; }

// ------------------------------------------------------------- utilities

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes)
    {
        if (bytes <= cap) return CPV_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) return fail(CPV_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
        cap = want;
        return CPV_OK;
    }
    template <class T> T* as() const { return (T*)p; }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct cpv_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_begin = nullptr, ev_raster0 = nullptr, ev_raster1 = nullptr, ev_end = nullptr;
    cpv_config cfg{};
    DevBuf vout, recs, counts, offsets, keys[2], vals[2], tile_start, hist, hist_base, sums, big;
    DevBuf arena, stats, traces;
    void* pinned = nullptr;
    size_t pinned_cap = 0;
    uint32_t* readback = nullptr;     // pinned scalar readbacks
    uint32_t last_tiles = 0;
    bool timing_valid = false;
    bool last_count_exact = false;
    float raster_ms_acc = 0.0f;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> raster_events;
    std::mutex lock;
};

extern "C" int cpv_context_create(int device_ordinal, cpv_context** out)
{
    if (!out) return fail(CPV_ERROR_INVALID, "out is null");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device_ordinal < 0 || device_ordinal >= n) return fail(CPV_ERROR_INVALID, "device %d of %d", device_ordinal, n);
    HIP_TRY(hipSetDevice(device_ordinal));
    cpv_context* c = new cpv_context();
    c->device = device_ordinal;
    c->cfg.tile_size = 16;
    c->cfg.waves_per_tile = 4;
    c->cfg.scheduler = CPV_SCHED_GRID;
    c->cfg.count_fragments = 1;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_begin) != hipSuccess || hipEventCreate(&c->ev_end) != hipSuccess ||
        hipEventCreate(&c->ev_raster0) != hipSuccess || hipEventCreate(&c->ev_raster1) != hipSuccess ||
        hipHostMalloc((void**)&c->readback, 256, hipHostMallocDefault) != hipSuccess) {
        delete c;
        return fail(CPV_ERROR_DEVICE, "stream/event/pinned allocation failed");
    }
    if (c->stats.ensure(256)) { delete c; return CPV_ERROR_OUT_OF_MEMORY; }
    *out = c;
    return CPV_OK;
}

extern "C" void cpv_context_destroy(cpv_context* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    DevBuf* bufs[] = {&c->vout, &c->recs, &c->counts, &c->offsets, &c->keys[0], &c->keys[1], &c->vals[0], &c->vals[1],
                      &c->tile_start, &c->big, &c->hist, &c->hist_base, &c->sums, &c->arena, &c->stats, &c->traces};
    for (DevBuf* b : bufs) b->release();
    if (c->pinned) (void)hipHostFree(c->pinned);
    if (c->readback) (void)hipHostFree(c->readback);
    (void)hipEventDestroy(c->ev_begin);
    (void)hipEventDestroy(c->ev_end);
    (void)hipEventDestroy(c->ev_raster0);
    (void)hipEventDestroy(c->ev_raster1);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int cpv_alloc(cpv_context* c, uint64_t bytes, uint64_t* address)
{
    if (!c || !address) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) return fail(CPV_ERROR_OUT_OF_MEMORY, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    *address = (uint64_t)(uintptr_t)p;
    return CPV_OK;
}

extern "C" int cpv_free(cpv_context* c, uint64_t address)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree((void*)(uintptr_t)address));
    return CPV_OK;
}

extern "C" int cpv_upload(cpv_context* c, uint64_t dst, const void* src, uint64_t bytes)
{
    if (!c || (!src && bytes)) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync((void*)(uintptr_t)dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPV_OK;
}

extern "C" int cpv_download(cpv_context* c, void* dst, uint64_t src, uint64_t bytes)
{
    if (!c || (!dst && bytes)) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(dst, (const void*)(uintptr_t)src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPV_OK;
}

__global__ void copy_u32x4_kernel(uint4* dst, const uint4* src, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

extern "C" int cpv_device_copy(cpv_context* c, uint64_t dst, uint64_t src, uint64_t bytes)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    if ((dst | src | bytes) & 15u) return fail(CPV_ERROR_INVALID, "device copy needs 16-byte aligned addresses and size");
    HIP_TRY(hipSetDevice(c->device));
    const uint64_t n = bytes / 16;
    if (n) {
        hipLaunchKernelGGL(copy_u32x4_kernel, dim3((uint32_t)std::min<uint64_t>((n + 255) / 256, 65536)), dim3(256), 0, c->stream,
                           (uint4*)(uintptr_t)dst, (const uint4*)(uintptr_t)src, n);
        HIP_TRY(hipGetLastError());
    }
    return CPV_OK;
}

extern "C" int cpv_synchronize(cpv_context* c)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CPV_OK;
}

static bool valid_tile_config(uint32_t t, uint32_t w)
{
    if (t == 8) return w == 1;
    if (t == 16 || t == 32) return w == 1 || w == 2 || w == 4;
    return false;
}

extern "C" int cpv_set_config(cpv_context* c, const cpv_config* cfg)
{
    if (!c || !cfg) return fail(CPV_ERROR_INVALID, "null argument");
    if (!valid_tile_config(cfg->tile_size, cfg->waves_per_tile))
        return fail(CPV_ERROR_INVALID, "unsupported tile %u x waves %u", cfg->tile_size, cfg->waves_per_tile);
    if (cfg->scheduler > CPV_SCHED_PREDICTED) return fail(CPV_ERROR_INVALID, "scheduler %u", cfg->scheduler);
    c->cfg = *cfg;
    return CPV_OK;
}

extern "C" int cpv_get_config(cpv_context* c, cpv_config* cfg)
{
    if (!c || !cfg) return fail(CPV_ERROR_INVALID, "null argument");
    *cfg = c->cfg;
    return CPV_OK;
}

// --------------------------------------------------------- kernel tables

namespace {

using RasterFn = void (*)(RasterArgs);
using VertexFn = void (*)(VertexArgs);

int tile_config_index(uint32_t t, uint32_t w)
{
    if (t == 8) return 0;
    if (t == 16) return w == 1 ? 1 : w == 2 ? 2 : 3;
    return w == 1 ? 4 : w == 2 ? 5 : 6;
}

template <class FS>
struct RasterTable {
    static const void* get(int i)
    {
        static const void* fns[7] = {
            (const void*)&raster_kernel<FS, 8, 1>,  (const void*)&raster_kernel<FS, 16, 1>,
            (const void*)&raster_kernel<FS, 16, 2>, (const void*)&raster_kernel<FS, 16, 4>,
            (const void*)&raster_kernel<FS, 32, 1>, (const void*)&raster_kernel<FS, 32, 2>,
            (const void*)&raster_kernel<FS, 32, 4>};
        return fns[i];
    }
};

struct Program {
    KernelRef vertex;
    KernelRef raster[7];
    int vs_outputs = 0;
    int fs_inputs = 0;
};

bool builtin_program(uint32_t vs, uint32_t fs, Program* p)
{
    switch (vs) {
    case CPV_VS_PASSTHROUGH: p->vertex.host_fn = (const void*)&vertex_kernel<VsPassthrough>; p->vs_outputs = VsPassthrough::kOutputs; break;
    case CPV_VS_MVP_COLOR: p->vertex.host_fn = (const void*)&vertex_kernel<VsMvpColor>; p->vs_outputs = VsMvpColor::kOutputs; break;
    case CPV_VS_MVP_LAMBERT: p->vertex.host_fn = (const void*)&vertex_kernel<VsMvpLambert>; p->vs_outputs = VsMvpLambert::kOutputs; break;
    default: return false;
    }
    switch (fs) {
    case CPV_FS_COLOR:
        for (int i = 0; i < 7; i++) p->raster[i].host_fn = RasterTable<FsColor>::get(i);
        p->fs_inputs = FsColor::kInputs;
        break;
    case CPV_FS_TEX_LAMBERT:
        for (int i = 0; i < 7; i++) p->raster[i].host_fn = RasterTable<FsTexLambert>::get(i);
        p->fs_inputs = FsTexLambert::kInputs;
        break;
    default: return false;
    }
    return true;
}

int resolve_program(const cpv_pipeline_desc& pd, Program* p)
{
    if (pd.program) {
        const cpv_program* prog = pd.program;
        p->vertex = prog->vertex;
        for (int i = 0; i < 7; i++) p->raster[i] = prog->raster[i];
        p->vs_outputs = prog->vs_outputs;
        p->fs_inputs = prog->fs_inputs;
        return CPV_OK;
    }
    if (!builtin_program(pd.vs_kind, pd.fs_kind, p))
        return fail(CPV_ERROR_UNSUPPORTED, "unknown built-in program vs=%u fs=%u", pd.vs_kind, pd.fs_kind);
    return CPV_OK;
}

hipError_t launch(const KernelRef& k, dim3 grid, dim3 block, void* args, size_t args_size, hipStream_t s)
{
    void* params[] = {args};
    if (k.host_fn) return hipLaunchKernel(k.host_fn, grid, block, params, 0, s);
    (void)args_size;
    return hipModuleLaunchKernel(k.module_fn, grid.x, grid.y, grid.z, block.x, block.y, block.z, 0, s, params, nullptr);
}

// ------------------------------------------------------ descriptor build

int build_fragment_state(const cpv_pipeline_desc& p, const cpv_pass_desc& pass, FragmentState* fs)
{
    std::memset(fs, 0, sizeof(*fs));
    fs->has_color = pass.color_count > 0 && p.color_count > 0;
    fs->has_ds = pass.has_depth_stencil != 0;
    if (fs->has_color && !cpv_format_desc(pass.color[0].format, &fs->color))
        return fail(CPV_ERROR_UNSUPPORTED, "colour format %u", pass.color[0].format);
    if (fs->has_ds && !cpv_format_desc(pass.depth_stencil.format, &fs->ds))
        return fail(CPV_ERROR_UNSUPPORTED, "depth/stencil format %u", pass.depth_stencil.format);
    if (fs->has_color && is_depth_stencil(fs->color)) return fail(CPV_ERROR_INVALID, "depth format used as colour");
    if (fs->has_ds && !is_depth_stencil(fs->ds)) return fail(CPV_ERROR_INVALID, "colour format used as depth");
    // The reference aborts (TODO_ERROR) on these (PipelineCompiler.cpp:1674-1700,
    // CommandBuffer.Draw.cpp:1602-1605, 1675-1678, :1232-1236).
    if (p.logic_op_enable) return fail(CPV_ERROR_UNSUPPORTED, "logic op");
    if (p.rasterizer_discard) return fail(CPV_ERROR_UNSUPPORTED, "rasterizer discard");
    if (p.polygon_mode != CPV_POLYGON_FILL) return fail(CPV_ERROR_UNSUPPORTED, "polygon mode %u", p.polygon_mode);
    if (p.depth_clamp && p.depth_test) return fail(CPV_ERROR_UNSUPPORTED, "depth clamp");
    if (fs->has_color && p.blend[0].color_write_mask != CPV_COLOR_RGBA) return fail(CPV_ERROR_UNSUPPORTED, "colour write mask");
    if (fs->has_color && p.blend[0].blend_enable && (fs->color.base == kBaseUInt || fs->color.base == kBaseSInt))
        return fail(CPV_ERROR_UNSUPPORTED, "blend on an integer attachment");
    fs->depth_test = p.depth_test != 0;
    fs->depth_write = p.depth_write != 0;
    fs->depth_compare = (int32_t)p.depth_compare;
    fs->depth_bounds_test = p.depth_bounds_test != 0;
    fs->stencil_test = p.stencil_test != 0;
    fs->min_depth_bounds = p.min_depth_bounds;
    fs->max_depth_bounds = p.max_depth_bounds;
    if (p.color_count) {
        const cpv_blend_attachment& b = p.blend[0];
        fs->blend_enable = b.blend_enable != 0;
        fs->src_color = (int32_t)b.src_color_factor; fs->dst_color = (int32_t)b.dst_color_factor;
        fs->color_op = (int32_t)b.color_op;
        fs->src_alpha = (int32_t)b.src_alpha_factor; fs->dst_alpha = (int32_t)b.dst_alpha_factor;
        fs->alpha_op = (int32_t)b.alpha_op;
        if (b.blend_enable && (b.src_color_factor > CPV_BLEND_SRC_ALPHA_SATURATE || b.dst_color_factor > CPV_BLEND_SRC_ALPHA_SATURATE ||
                               b.src_alpha_factor > CPV_BLEND_SRC_ALPHA_SATURATE || b.dst_alpha_factor > CPV_BLEND_SRC_ALPHA_SATURATE ||
                               b.color_op > CPV_BLEND_OP_MAX || b.alpha_op > CPV_BLEND_OP_MAX))
            return fail(CPV_ERROR_UNSUPPORTED, "blend factor/op");
    }
    for (int i = 0; i < 4; i++) fs->blend_constants[i] = p.blend_constants[i];
    const cpv_stencil_state* st[2] = {&p.front, &p.back};
    int32_t* dst[2] = {fs->st_front, fs->st_back};
    for (int s = 0; s < 2; s++) {
        dst[s][0] = (int32_t)st[s]->fail_op; dst[s][1] = (int32_t)st[s]->pass_op;
        dst[s][2] = (int32_t)st[s]->depth_fail_op; dst[s][3] = (int32_t)st[s]->compare_op;
        dst[s][4] = (int32_t)st[s]->compare_mask; dst[s][5] = (int32_t)st[s]->write_mask;
        dst[s][6] = (int32_t)st[s]->reference;
    }
    fs->origin_upper_left = p.origin_upper_left != 0;
    return CPV_OK;
}

int build_texture(const cpv_texture_desc& t, TextureDesc* o)
{
    std::memset(o, 0, sizeof(*o));
    if (!t.address) return CPV_OK;
    if (!cpv_format_desc(t.format, &o->fmt) || is_depth_stencil(o->fmt)) return fail(CPV_ERROR_UNSUPPORTED, "texture format %u", t.format);
    if (t.levels == 0 || t.levels > 16 || t.base_level + t.level_count > t.levels || t.level_count == 0)
        return fail(CPV_ERROR_INVALID, "texture levels");
    o->address = t.address;
    o->level_count = t.level_count;
    // GetNormalImageSize (CPVulkanBase/Formats.cpp:455-482)
    uint64_t off = 0;
    uint32_t w = t.width, h = t.height;
    for (uint32_t i = 0; i < t.levels; i++) {
        if (i >= t.base_level && i < t.base_level + t.level_count) {
            const uint32_t j = i - t.base_level;
            o->level_offset[j] = off;
            o->level_w[j] = w;
            o->level_h[j] = h;
        }
        off += (uint64_t)o->fmt.total * w * h;
        w = std::max(w / 2, 1u);
        h = std::max(h / 2, 1u);
    }
    const cpv_sampler_desc& s = t.sampler;
    if (s.unnormalized) return fail(CPV_ERROR_UNSUPPORTED, "unnormalised coordinates");
    o->sampler.mag_filter = (int32_t)s.mag_filter;
    o->sampler.min_filter = (int32_t)s.min_filter;
    o->sampler.mipmap_mode = (int32_t)s.mipmap_mode;
    o->sampler.address_u = (int32_t)s.address_u;
    o->sampler.address_v = (int32_t)s.address_v;
    o->sampler.mip_lod_bias = s.mip_lod_bias;
    o->sampler.min_lod = s.min_lod;
    o->sampler.max_lod = s.max_lod;
    o->sampler.border_color = (int32_t)s.border_color;
    return CPV_OK;
}

int build_env(const cpv_pipeline_desc& p, const cpv_draw_desc& d, DrawEnv* e)
{
    std::memset(e, 0, sizeof(*e));
    for (uint32_t i = 0; i < p.attribute_count && i < CPV_MAX_VERTEX_ATTRIBUTES; i++) {
        const cpv_vertex_attribute_desc& a = p.attributes[i];
        if (a.location >= 16) return fail(CPV_ERROR_INVALID, "attribute location %u", a.location);
        AttribDesc& o = e->attr[a.location];
        if (!cpv_format_desc(a.format, &o.fmt)) return fail(CPV_ERROR_UNSUPPORTED, "attribute format %u", a.format);
        o.binding = a.binding;
        o.offset = a.offset;
        o.valid = 1;
    }
    for (uint32_t i = 0; i < p.binding_count && i < CPV_MAX_VERTEX_BINDINGS; i++) {
        const cpv_vertex_binding_desc& b = p.bindings[i];
        if (b.binding >= 16) return fail(CPV_ERROR_INVALID, "binding %u", b.binding);
        BindingDesc& o = e->bind[b.binding];
        o.address = d.vertex_address[b.binding];
        o.bytes = d.vertex_bytes[b.binding];
        o.stride = b.stride;
        o.input_rate = b.input_rate;
    }
    for (int i = 0; i < 4; i++) e->uniform[i] = d.uniform_address[i];
    for (int i = 0; i < 4; i++) {
        int rc = build_texture(d.textures[i], &e->tex[i]);
        if (rc) return rc;
    }
    return CPV_OK;
}

int build_target(const cpv_pass_desc& pass, bool first_segment, int tile, uint32_t tile_base, TargetParams* t)
{
    std::memset(t, 0, sizeof(*t));
    if (pass.width == 0 || pass.height == 0 || pass.width > 16384 || pass.height > 16384)
        return fail(CPV_ERROR_INVALID, "pass extent %ux%u", pass.width, pass.height);
    t->width = pass.width;
    t->height = pass.height;
    t->tiles_x = (pass.width + tile - 1) / tile;
    t->tiles_y = (pass.height + tile - 1) / tile;
    t->tile_base = tile_base;
    t->has_color = pass.color_count > 0 && pass.color[0].address != 0;
    t->has_ds = pass.has_depth_stencil && pass.depth_stencil.address != 0;
    FormatDesc f;
    if (t->has_color) {
        if (!cpv_format_desc(pass.color[0].format, &f)) return fail(CPV_ERROR_UNSUPPORTED, "colour format %u", pass.color[0].format);
        t->color_fmt = f;
        t->color_addr = pass.color[0].address;
        t->color_bytes = f.total;
        // DONT_CARE keeps memory as is in the reference, which is LOAD here.
        t->color_load = (first_segment && pass.color[0].load_op == CPV_LOAD_OP_CLEAR) ? 1 : 0;
        for (int i = 0; i < 4; i++) t->clear_color[i] = pass.color[0].clear.u32[i];
    }
    if (t->has_ds) {
        if (!cpv_format_desc(pass.depth_stencil.format, &f)) return fail(CPV_ERROR_UNSUPPORTED, "depth format %u", pass.depth_stencil.format);
        t->ds_fmt = f;
        t->ds_addr = pass.depth_stencil.address;
        t->ds_bytes = f.total;
        t->ds_load = (first_segment && pass.depth_stencil.load_op == CPV_LOAD_OP_CLEAR) ? 1 : 0;
        std::memcpy(&t->clear_ds[0], &pass.depth_stencil.clear.ds.depth, 4);
        t->clear_ds[1] = pass.depth_stencil.clear.ds.stencil;
    }
    return CPV_OK;
}

uint32_t prim_count_of(uint32_t topology, uint32_t n)
{
    if (topology == CPV_TOPOLOGY_TRIANGLE_LIST) return n / 3;
    if (topology == CPV_TOPOLOGY_TRIANGLE_STRIP || topology == CPV_TOPOLOGY_TRIANGLE_FAN) return n > 2 ? n - 2 : 0;
    return 0;
}

}  // namespace

// ----------------------------------------------------- launch sequence

// Host staging of one launch's descriptors, copied to the device arena in one
// transfer.
struct Staging {
    std::vector<uint8_t> bytes;
    size_t put(const void* p, size_t n, size_t align = 256)
    {
        size_t at = (bytes.size() + align - 1) & ~(align - 1);
        bytes.resize(at + n);
        std::memcpy(bytes.data() + at, p, n);
        return at;
    }
};

struct LaunchDraw {
    const cpv_pipeline_desc* pipe;
    const cpv_draw_desc* draw;
    uint32_t instance;
};

static int scan_u32(cpv_context* c, const uint32_t* in, uint32_t n, uint32_t* out, bool write_total)
{
    if (n == 0) {
        if (write_total) HIP_TRY(hipMemsetAsync(out, 0, 4, c->stream));
        return CPV_OK;
    }
    const uint32_t blocks = (n + kScanTile - 1) / kScanTile;
    if (c->sums.ensure((size_t)blocks * 4)) return CPV_ERROR_OUT_OF_MEMORY;
    uint32_t* sums = c->sums.as<uint32_t>();
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(blocks), dim3(kScanThreads), 0, c->stream, in, n, sums);
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, c->stream, sums, blocks);
    hipLaunchKernelGGL(scan_down_kernel, dim3(blocks), dim3(kScanThreads), 0, c->stream, in, n, (const uint32_t*)sums, out,
                       write_total ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return CPV_OK;
}

static void fill_u32(cpv_context* c, void* p, uint32_t v, uint64_t n)
{
    if (!n) return;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, c->stream, (uint32_t*)p, v, n);
}

// Runs one launch: draws (all with the same program) over targets.
static int run_launch(cpv_context* c, const Program& prog, const std::vector<TargetParams>& targets,
                      const std::vector<LaunchDraw>& lds, const std::vector<cpv_pass_desc>& passes,
                      const std::vector<uint32_t>& draw_target, bool record_timing)
{
    const int tile = (int)c->cfg.tile_size;
    const int tile_shift = tile == 8 ? 3 : tile == 16 ? 4 : 5;
    const uint32_t stride = 4u + (uint32_t)prog.vs_outputs;
    if (prog.fs_inputs > prog.vs_outputs) return fail(CPV_ERROR_INVALID, "fragment inputs exceed vertex outputs");

    std::vector<DrawDesc> dd(lds.size());
    std::vector<DrawEnv> envs(lds.size());
    std::vector<FragmentState> fstates(lds.size());
    std::vector<uint32_t> vblock, sblock;
    uint32_t vout_total = 0, prim_total = 0;
    for (size_t i = 0; i < lds.size(); i++) {
        const cpv_pipeline_desc& p = *lds[i].pipe;
        const cpv_draw_desc& d = *lds[i].draw;
        int rc = build_env(p, d, &envs[i]);
        if (rc) return rc;
        rc = build_fragment_state(p, passes[draw_target[i]], &fstates[i]);
        if (rc) return rc;
        if (p.topology < CPV_TOPOLOGY_TRIANGLE_LIST || p.topology > CPV_TOPOLOGY_TRIANGLE_FAN)
            return fail(CPV_ERROR_UNSUPPORTED, "topology %u (lines and points are not on this path)", p.topology);
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
        DrawDesc& o = dd[i];
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
        o.vblock_base = (uint32_t)vblock.size();
        o.sblock_base = (uint32_t)sblock.size();
        for (uint32_t b = 0; b < (nverts + 255) / 256; b++) vblock.push_back((uint32_t)i);
        for (uint32_t b = 0; b < (o.prim_count + 255) / 256; b++) sblock.push_back((uint32_t)i);
        vout_total += nverts;
        prim_total += o.prim_count;
    }
    if (prim_total >= (1u << 28)) return fail(CPV_ERROR_UNSUPPORTED, "more than 2^28 primitives in one launch");
    uint32_t tiles = 0;
    for (const TargetParams& t : targets) tiles = std::max(tiles, t.tile_base + t.tiles_x * t.tiles_y);

    // descriptors -> device arena (the block tables are never empty)
    if (vblock.empty()) vblock.push_back(0);
    if (sblock.empty()) sblock.push_back(0);
    const bool any_vertices = vout_total > 0, any_prims = prim_total > 0;
    Staging st;
    const size_t o_env = st.put(envs.data(), envs.size() * sizeof(DrawEnv));
    const size_t o_dd = st.put(dd.data(), dd.size() * sizeof(DrawDesc));
    const size_t o_fs = st.put(fstates.data(), fstates.size() * sizeof(FragmentState));
    const size_t o_tp = st.put(targets.data(), targets.size() * sizeof(TargetParams));
    const size_t o_vb = st.put(vblock.data(), std::max<size_t>(vblock.size(), 1) * 4);
    const size_t o_sb = st.put(sblock.data(), std::max<size_t>(sblock.size(), 1) * 4);
    if (c->arena.ensure(st.bytes.size())) return CPV_ERROR_OUT_OF_MEMORY;
    uint8_t* arena = c->arena.as<uint8_t>();
    for (size_t i = 0; i < dd.size(); i++) dd[i].env = (uint64_t)(uintptr_t)(arena + o_env + i * sizeof(DrawEnv));
    std::memcpy(st.bytes.data() + o_dd, dd.data(), dd.size() * sizeof(DrawDesc));
    if (st.bytes.size() > c->pinned_cap) {
        if (c->pinned) (void)hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_cap = 0;
        HIP_TRY(hipHostMalloc(&c->pinned, st.bytes.size() * 2, hipHostMallocDefault));
        c->pinned_cap = st.bytes.size() * 2;
    }
    // The previous launch may still read the pinned staging: wait for it.
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(c->pinned, st.bytes.data(), st.bytes.size());
    HIP_TRY(hipMemcpyAsync(arena, c->pinned, st.bytes.size(), hipMemcpyHostToDevice, c->stream));
    const DrawDesc* d_dd = (const DrawDesc*)(arena + o_dd);
    const FragmentState* d_fs = (const FragmentState*)(arena + o_fs);
    const TargetParams* d_tp = (const TargetParams*)(arena + o_tp);
    const uint32_t* d_vb = (const uint32_t*)(arena + o_vb);
    const uint32_t* d_sb = (const uint32_t*)(arena + o_sb);

    // buffers
    if (c->vout.ensure((size_t)std::max(vout_total, 1u) * stride * 4) || c->recs.ensure((size_t)std::max(prim_total, 1u) * sizeof(TriSetup)) ||
        c->counts.ensure((size_t)std::max(prim_total, 1u) * 4) || c->offsets.ensure(((size_t)prim_total + 1) * 4) ||
        c->tile_start.ensure(((size_t)tiles + 1) * 4) ||
        c->hist.ensure((size_t)256 * kSortSegments * 4) || c->hist_base.ensure(((size_t)256 * kSortSegments + 1) * 4))
        return CPV_ERROR_OUT_OF_MEMORY;

    // 1. vertex stage
    if (any_vertices) {
        VertexArgs va{d_dd, d_vb, c->vout.as<float>(), stride, 0};
        HIP_TRY(launch(prog.vertex, dim3((uint32_t)vblock.size()), dim3(256), &va, sizeof(va), c->stream));
    }
    // 2. triangle setup
    if (any_prims) {
        SetupArgs sa{d_dd, d_sb, d_tp, c->vout.as<float>(), c->recs.as<TriSetup>(), c->counts.as<uint32_t>(), stride, tile_shift};
        hipLaunchKernelGGL(setup_kernel, dim3((uint32_t)sblock.size()), dim3(256), 0, c->stream, sa);
        HIP_TRY(hipGetLastError());
    }
    // 3. pair offsets and their total
    int rc = scan_u32(c, c->counts.as<uint32_t>(), prim_total, c->offsets.as<uint32_t>(), true);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->readback, c->offsets.as<uint32_t>() + prim_total, 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t pairs = c->readback[0];
    if (c->keys[0].ensure((size_t)std::max(pairs, 1u) * 4) || c->keys[1].ensure((size_t)std::max(pairs, 1u) * 4) ||
        c->vals[0].ensure((size_t)std::max(pairs, 1u) * 4) || c->vals[1].ensure((size_t)std::max(pairs, 1u) * 4))
        return CPV_ERROR_OUT_OF_MEMORY;
    // 4. emit pairs in primitive order
    if (prim_total) {
        if (c->big.ensure(((size_t)prim_total + 1) * 4)) return CPV_ERROR_OUT_OF_MEMORY;
        fill_u32(c, c->big.p, 0u, 1);
        EmitArgs ea{c->recs.as<TriSetup>(), c->counts.as<uint32_t>(), c->offsets.as<uint32_t>(), d_dd, d_tp,
                    c->keys[0].as<uint32_t>(), c->vals[0].as<uint32_t>(), c->big.as<uint32_t>(), prim_total, tile_shift};
        const uint32_t blocks = std::min<uint32_t>((prim_total + 255) / 256, 8192);
        hipLaunchKernelGGL(emit_kernel, dim3(blocks), dim3(256), 0, c->stream, ea);
        hipLaunchKernelGGL(emit_big_kernel, dim3(1024), dim3(256), 0, c->stream, ea);
        HIP_TRY(hipGetLastError());
    }
    // 5. stable radix sort of the pairs by tile
    int src = 0;
    if (pairs > 1 && tiles > 1) {
        int bits = 0;
        while ((1u << bits) < tiles) bits++;
        uint32_t* d_n = c->counts.as<uint32_t>();   // reuse: counts no longer needed after emit
        fill_u32(c, d_n, pairs, 1);
        for (int shift = 0; shift < bits; shift += 8) {
            hipLaunchKernelGGL(radix_hist_kernel, dim3(kSortSegments / kSortWaves), dim3(64 * kSortWaves), 0, c->stream,
                               (const uint32_t*)c->keys[src].p, (const uint32_t*)d_n, shift, c->hist.as<uint32_t>());
            rc = scan_u32(c, c->hist.as<uint32_t>(), 256u * kSortSegments, c->hist_base.as<uint32_t>(), false);
            if (rc) return rc;
            hipLaunchKernelGGL(radix_scatter_kernel, dim3(kSortSegments / kSortWaves), dim3(64 * kSortWaves), 0, c->stream,
                               (const uint32_t*)c->keys[src].p, (const uint32_t*)c->vals[src].p, c->keys[1 - src].as<uint32_t>(),
                               c->vals[1 - src].as<uint32_t>(), (const uint32_t*)d_n, shift, (const uint32_t*)c->hist_base.p);
            HIP_TRY(hipGetLastError());
            src = 1 - src;
        }
    }
    // tile ranges from the sorted keys
    hipLaunchKernelGGL(tile_ranges_kernel, dim3(pairs / 256 + 1), dim3(256), 0, c->stream, (const uint32_t*)c->keys[src].p, pairs,
                       tiles, c->tile_start.as<uint32_t>());
    HIP_TRY(hipGetLastError());
    // 6. raster
    const int ci = tile_config_index(c->cfg.tile_size, c->cfg.waves_per_tile);
    const KernelRef& rk = prog.raster[ci];
    RasterArgs ra{};
    ra.recs = c->recs.as<TriSetup>();
    ra.tile_prims = c->vals[src].as<uint32_t>();
    ra.tile_start = c->tile_start.as<uint32_t>();
    ra.targets = d_tp;
    ra.draws = d_dd;
    ra.fstates = d_fs;
    ra.vout = c->vout.as<float>();
    ra.order = nullptr;
    ra.queue = c->stats.as<uint32_t>() + 8;
    ra.fragments = (unsigned long long*)c->stats.p;
    ra.trace_cycles = nullptr;
    ra.vout_stride = stride;
    ra.tiles = tiles;
    ra.target_count = (uint32_t)targets.size();
    ra.persistent = 0;
    // Hierarchical depth culling is exact when only the depth test observes
    // a fragment (no stencil, no depth bounds), every draw uses the same
    // compare op and depth format, and nobody counts covered fragments.
    ra.hiz_op = -1;
    ra.hiz_depth_kind = 32;
    if (!c->cfg.count_fragments && !fstates.empty()) {
        bool ok = true;
        for (const FragmentState& f : fstates) {
            ok = ok && f.depth_test && f.has_ds && f.ds.depth_off >= 0 && f.ds.vk_format != CPV_FORMAT_S8_UINT &&
                 !f.depth_bounds_test && !f.stencil_test && f.depth_compare == fstates[0].depth_compare &&
                 f.ds.depth_kind == fstates[0].ds.depth_kind;
        }
        for (const TargetParams& t : targets) ok = ok && t.has_ds && t.ds_fmt.depth_kind == fstates[0].ds.depth_kind;
        if (ok) {
            ra.hiz_op = fstates[0].depth_compare;
            ra.hiz_depth_kind = fstates[0].ds.depth_kind;
        }
    }
    c->last_count_exact = c->cfg.count_fragments != 0;
    if (c->cfg.record_costs) {
        if (c->traces.ensure((size_t)std::max(tiles, 1u) * c->cfg.waves_per_tile * 4)) return CPV_ERROR_OUT_OF_MEMORY;
        ra.trace_cycles = c->traces.as<float>();
    }
    if (record_timing) HIP_TRY(hipEventRecord(c->ev_raster0, c->stream));
    HIP_TRY(launch(rk, dim3(tiles), dim3(64 * c->cfg.waves_per_tile), &ra, sizeof(ra), c->stream));
    if (record_timing) HIP_TRY(hipEventRecord(c->ev_raster1, c->stream));
    c->last_tiles = tiles;
    return CPV_OK;
}

static bool same_program(const cpv_pipeline_desc& a, const cpv_pipeline_desc& b)
{
    if (a.program || b.program) return a.program == b.program;
    return a.vs_kind == b.vs_kind && a.fs_kind == b.fs_kind;
}

extern "C" int cpv_render_pass(cpv_context* c, const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                               const cpv_draw_desc* draws, const uint32_t* draw_pipeline, uint32_t draw_count)
{
    if (!c || !pass || (draw_count && (!pipelines || !draws))) return fail(CPV_ERROR_INVALID, "null argument");
    if (pass->color_count > CPV_MAX_COLOR_ATTACHMENTS) return fail(CPV_ERROR_UNSUPPORTED, "%u colour attachments", pass->color_count);
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    fill_u32(c, c->stats.p, 0u, 64);
    HIP_TRY(hipEventRecord(c->ev_begin, c->stream));
    const int tile = (int)c->cfg.tile_size;
    std::vector<cpv_pass_desc> passes(1, *pass);
    std::vector<uint32_t> order;
    bool first = true;
    uint32_t i = 0;
    do {
        // a run of draws sharing one shader program
        std::vector<LaunchDraw> lds;
        std::vector<uint32_t> dt;
        Program prog;
        if (draw_count == 0) {
            // clear-only pass: an empty launch still initialises every tile
            if (!builtin_program(CPV_VS_PASSTHROUGH, CPV_FS_COLOR, &prog)) return fail(CPV_ERROR_DEVICE, "built-in table");
        } else {
            const uint32_t pi0 = draw_pipeline ? draw_pipeline[i] : 0;
            int rc = resolve_program(pipelines[pi0], &prog);
            if (rc) return rc;
            uint32_t j = i;
            while (j < draw_count) {
                const uint32_t pj = draw_pipeline ? draw_pipeline[j] : 0;
                if (!same_program(pipelines[pj], pipelines[pi0])) break;
                for (uint32_t inst = 0; inst < draws[j].instance_count; inst++) {
                    lds.push_back({&pipelines[pj], &draws[j], inst});
                    dt.push_back(0);
                }
                j++;
            }
            i = j;
        }
        std::vector<TargetParams> targets(1);
        int rc = build_target(*pass, first, tile, 0, &targets[0]);
        if (rc) return rc;
        rc = run_launch(c, prog, targets, lds, passes, dt, true);
        if (rc) return rc;
        first = false;
    } while (i < draw_count);
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    c->timing_valid = true;
    return CPV_OK;
}

extern "C" int cpv_render_jobs(cpv_context* c, const cpv_job_desc* jobs, uint32_t job_count)
{
    if (!c || (job_count && !jobs)) return fail(CPV_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> guard(c->lock);
    HIP_TRY(hipSetDevice(c->device));
    fill_u32(c, c->stats.p, 0u, 64);
    HIP_TRY(hipEventRecord(c->ev_begin, c->stream));
    const int tile = (int)c->cfg.tile_size;
    // jobs are grouped by shader program; each group is one launch of every stage
    std::vector<bool> done(job_count, false);
    for (uint32_t g = 0; g < job_count; g++) {
        if (done[g]) continue;
        Program prog;
        int rc = resolve_program(jobs[g].pipeline, &prog);
        if (rc) return rc;
        std::vector<TargetParams> targets;
        std::vector<cpv_pass_desc> passes;
        std::vector<LaunchDraw> lds;
        std::vector<uint32_t> dt;
        uint32_t tile_base = 0;
        for (uint32_t j = g; j < job_count; j++) {
            if (done[j] || !same_program(jobs[j].pipeline, jobs[g].pipeline)) continue;
            done[j] = true;
            TargetParams t;
            rc = build_target(jobs[j].pass, true, tile, tile_base, &t);
            if (rc) return rc;
            tile_base += t.tiles_x * t.tiles_y;
            const uint32_t ti = (uint32_t)targets.size();
            targets.push_back(t);
            passes.push_back(jobs[j].pass);
            for (uint32_t inst = 0; inst < jobs[j].draw.instance_count; inst++) {
                lds.push_back({&jobs[j].pipeline, &jobs[j].draw, inst});
                dt.push_back(ti);
            }
        }
        rc = run_launch(c, prog, targets, lds, passes, dt, true);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(c->ev_end, c->stream));
    c->timing_valid = true;
    return CPV_OK;
}

extern "C" int cpv_last_fragment_count(cpv_context* c, uint64_t* fragments)
{
    if (!c || !fragments) return fail(CPV_ERROR_INVALID, "null argument");
    if (!c->last_count_exact) return fail(CPV_ERROR_INVALID, "the last render call ran with count_fragments off");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(c->readback + 16, c->stats.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::memcpy(fragments, c->readback + 16, 8);
    return CPV_OK;
}

extern "C" int cpv_last_timing(cpv_context* c, float* raster_ms, float* total_ms)
{
    if (!c) return fail(CPV_ERROR_INVALID, "null context");
    if (!c->timing_valid) return fail(CPV_ERROR_INVALID, "no render call recorded");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipEventSynchronize(c->ev_end));
    if (raster_ms) HIP_TRY(hipEventElapsedTime(raster_ms, c->ev_raster0, c->ev_raster1));
    if (total_ms) HIP_TRY(hipEventElapsedTime(total_ms, c->ev_begin, c->ev_end));
    return CPV_OK;
}

extern "C" int cpv_get_tile_traces(cpv_context* c, cpv_tile_trace* traces, uint32_t capacity, uint32_t* count)
{
    if (!c || !count) return fail(CPV_ERROR_INVALID, "null argument");
    *count = 0;
    (void)traces;
    (void)capacity;
    return fail(CPV_ERROR_UNSUPPORTED, "tile traces are read through the scheduler module");
}

extern "C" int cpv_abi_sizes(uint32_t* out, uint32_t n)
{
    const uint32_t sizes[] = {(uint32_t)sizeof(cpv_pipeline_desc), (uint32_t)sizeof(cpv_draw_desc), (uint32_t)sizeof(cpv_pass_desc),
                              (uint32_t)sizeof(cpv_attachment_desc), (uint32_t)sizeof(cpv_texture_desc), (uint32_t)sizeof(cpv_sampler_desc),
                              (uint32_t)sizeof(cpv_job_desc), (uint32_t)sizeof(cpv_config), (uint32_t)sizeof(cpv_tile_trace)};
    const uint32_t m = (uint32_t)(sizeof(sizes) / sizeof(sizes[0]));
    for (uint32_t i = 0; i < n && i < m; i++) out[i] = sizes[i];
    return (int)m;
}

// ------------------------------------------------------- SPIR-V programs

int cpv_spirv_compile(cpv_context* c, const uint32_t* vs, size_t vs_n, const char* vs_entry, const uint32_t* fs, size_t fs_n,
                      const char* fs_entry, cpv_program** out, std::string* err);

extern "C" int cpv_program_create_spirv(cpv_context* c, const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry,
                                        const uint32_t* fs_words, size_t fs_word_count, const char* fs_entry, cpv_program** out)
{
    if (!c || !vs_words || !fs_words || !out) return fail(CPV_ERROR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    std::string err;
    int rc = cpv_spirv_compile(c, vs_words, vs_word_count, vs_entry, fs_words, fs_word_count, fs_entry, out, &err);
    if (rc) return fail(rc, "%s", err.c_str());
    return CPV_OK;
}

extern "C" void cpv_program_destroy(cpv_context* c, cpv_program* p)
{
    if (!p) return;
    if (c) (void)hipSetDevice(c->device);
    if (p->module) (void)hipModuleUnload(p->module);
    delete p;
}

extern "C" const char* cpv_program_source(cpv_program* p) { return p ? p->source.c_str() : ""; }
